// Learned-seed network: parameters, BF16 tile packing, pair features,
// standalone inference, and training (forward + backward per wave on the
// matrix cores, deterministic per-block gradient partials, reduction, Adam).
//
// The network and its training have no counterpart in the reference; the
// training pairs come from the plain method (converged Newton solutions).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "mlp.hpp"
#include "rtn_fast.hpp"
#include "rtn_internal.hpp"

using namespace rtn;
using namespace rtn::mlp;

struct rtn_mlp {
  DevBuf<float> w, m, v, grad;  // FP32 master parameters, Adam moments, gradient
  DevBuf<bf16x8> tiles;         // NTILE x 64 lanes
  DevBuf<float> bias;           // NBIAS
  DevBuf<float> partial;        // per-block gradient partials
  DevBuf<float> loss;           // [0] running sum of squared errors
  int64_t partial_blocks = 0;
  int64_t step = 0;
  int kind = FEAT_PROJECTED;    // input feature kind (RTN_MLP_FEAT_*)
};

namespace {

constexpr int FWD_WAVES = 4;  // standalone inference: waves per block

__device__ __forceinline__ float wave_sum(float x) {
#pragma unroll
  for (int s = 32; s > 0; s >>= 1) x += __shfl_xor(x, s);
  return x;
}

// Master parameters -> BF16 operand tiles and FP32 bias seeds.
__global__ void k_pack(const float* __restrict__ w, bf16x8* __restrict__ tiles, float* __restrict__ bias) {
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  if (gid < NBIAS) {
    float b;
    if (gid < BB2) b = w[OFF_B1 + gid];
    else if (gid < BB3) b = w[OFF_B2 + gid - BB2];
    else if (gid < BB4) b = w[OFF_B3 + gid - BB3];
    else b = gid - BB4 < NOUT ? w[OFF_B4 + gid - BB4] : 0.0f;
    bias[gid] = b;
  }
  if (gid >= NTILE * 64) return;
  const int t = gid / 64, lane = gid % 64, m = lane & 15, g = lane >> 4;
  bf16x8 r;
#pragma unroll
  for (int i = 0; i < 8; i++) {
    float x;
    if (t < T_L2) {  // layer 1, row block t, natural k order (features)
      const int k = 8 * g + i;
      x = k < NIN ? w[OFF_W1 + (16 * t + m) * NIN + k] : 0.0f;
    } else if (t < T_L4) {  // layers 2, 3: k in the permuted slot order
      const int base = t < T_L3 ? T_L2 : T_L3, off = t < T_L3 ? OFF_W2 : OFF_W3;
      const int ob = (t - base) >> 1, kb = (t - base) & 1;
      x = w[off + (16 * ob + m) * NH + 32 * kb + perm_unit(g, i)];
    } else if (t < T_T2) {  // output layer, rows 0..1 real
      const int kb = t - T_L4;
      x = m < NOUT ? w[OFF_W4 + m * NH + 32 * kb + perm_unit(g, i)] : 0.0f;
    } else {  // transposed W2 / W3 for the backward pass, natural k order
      const int base = t < T_T3 ? T_T2 : T_T3, off = t < T_T3 ? OFF_W2 : OFF_W3;
      const int ob = (t - base) >> 1, kb = (t - base) & 1;
      x = w[off + (32 * kb + 8 * g + i) * NH + 16 * ob + m];
    }
    r[i] = (__bf16)x;
  }
  tiles[gid] = r;
}

__global__ void k_features(const float* __restrict__ bounds, const float4* __restrict__ cp,
                           const int* __restrict__ cp_off, const int* __restrict__ pq, int np, const float* ox,
                           const float* oy, const float* oz,
                           const float* dx, const float* dy, const float* dz, int64_t n_rays,
                           const int* __restrict__ pair_ray, const int* __restrict__ pair_patch, int64_t n,
                           int kind, float* __restrict__ feats, int* __restrict__ bad) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int r = pair_ray[i], p = pair_patch[i];
  if (r < 0 || r >= n_rays || p < 0 || p >= np) {
    atomicExch(bad, 1);
    return;
  }
  float f[NIN];
  const int v = pq[p];
  const rtn::fast::f3 o{ox[r], oy[r], oz[r]}, d{dx[r], dy[r], dz[r]};
  const rtn::fast::Plane2 rp = rtn::fast::planes(o, d);
  kind_features(kind, o.x, o.y, o.z, d.x, d.y, d.z, rp.n1.x, rp.n1.y, rp.n1.z, rp.n2.x, rp.n2.y, rp.n2.z,
                cp + cp_off[p], v & 0xff, (v >> 8) & 0xff, bounds + 6 * (int64_t)p, f);
#pragma unroll
  for (int k = 0; k < NIN; k++) feats[i * NIN + k] = f[k];
}

// Standalone inference: one wave per 64 pairs, the same per-wave routine as
// the trace kernel's learned seed.
__global__ void __launch_bounds__(64 * FWD_WAVES)
k_forward(const bf16x8* __restrict__ tiles, const float* __restrict__ bias, const float* __restrict__ feats,
          const int* __restrict__ valid, int64_t n, float* __restrict__ uv, float* __restrict__ hit) {
  __shared__ bf16x8 s_x[FWD_WAVES][64][4];
  __shared__ float4 s_logit[FWD_WAVES][64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int64_t i = ((int64_t)blockIdx.x * FWD_WAVES + wv) * 64 + lane;
  const bool active = i < n && (!valid || valid[i] != 0);
  float f[NIN];
#pragma unroll
  for (int k = 0; k < NIN; k++) f[k] = active ? feats[i * NIN + k] : 0.0f;
  float u = 0.f, v = 0.f, ph = 0.f;
  seed_wave(tiles, bias, active, f, u, v, ph, s_x[wv], s_logit[wv], lane);
  if (active) {
    uv[2 * i] = u;
    uv[2 * i + 1] = v;
    if (hit) hit[i] = ph;
  }
}

// Row pitch (BF16 values) of the pair-major gradient copy: 16-byte aligned
// rows that do not all start in the same LDS bank.
constexpr int DZP = 72;

// One training wave: forward (keeping pre-activations) and backward for 64
// pairs; writes this block's full gradient partial (NPARAM floats) and its
// loss sum.  A target is (u, v, h), h = 1 for a hit (converged solve), 0 for
// a miss.  Loss = mean over the batch of h |y_uv - (u,v)|^2 + lambda BCE(y_h, h),
// lambda = hit_weight.
__global__ void __launch_bounds__(64)
k_train(const bf16x8* __restrict__ tiles, const float* __restrict__ bias, const float* __restrict__ w,
        const float* __restrict__ feats, const float* __restrict__ targets, const int* __restrict__ perm, int64_t base,
        int64_t end, float inv_b, float hit_weight, float* __restrict__ partial, float* __restrict__ loss_part) {
  __shared__ bf16x8 s_x[64][4];
  __shared__ float4 s_logit[64];
  __shared__ float s_z[3 * NH * 64];  // pre-activations z1 | z2 | z3; z3's slice later holds dh
  __shared__ __bf16 s_dzu[NH * 64];   // [unit][pair]
  __shared__ __bf16 s_dzp[64 * DZP];  // [pair][unit]
  const int lane = threadIdx.x, m = lane & 15, g = lane >> 4;
  const int64_t pos = base + (int64_t)blockIdx.x * 64 + lane;
  const bool active = pos < end;
  const int64_t id = active ? (perm ? (int64_t)perm[pos] : pos) : 0;
  float f[NIN];
#pragma unroll
  for (int k = 0; k < NIN; k++) f[k] = active ? feats[id * NIN + k] : 0.0f;
  const float t0 = active ? targets[3 * id] : 0.f, t1 = active ? targets[3 * id + 1] : 0.f;
  const float th = active ? targets[3 * id + 2] : 0.f;
  store_row(s_x, lane, f, !active);
  wave_lds_sync();
  forward_wave<true>(tiles, bias, s_x, 4, lane, s_logit, s_z);
  wave_lds_sync();
  float* P = partial + (int64_t)blockIdx.x * NPARAM;

  // output layer: sigmoids; squared (u,v) error on hits, cross-entropy of the
  // hit probability (d BCE / d logit = y_h - h)
  const float4 lg = s_logit[lane];
  const float y0 = fast_sigmoid(lg.x), y1 = fast_sigmoid(lg.y), yh = fast_sigmoid(lg.z);
  const bool hit = active && th > 0.5f;
  const float e0 = hit ? y0 - t0 : 0.f, e1 = hit ? y1 - t1 : 0.f;
  const float dz40 = 2.0f * e0 * y0 * (1.0f - y0) * inv_b;
  const float dz41 = 2.0f * e1 * y1 * (1.0f - y1) * inv_b;
  const float dz42 = active ? hit_weight * (yh - th) * inv_b : 0.f;
  // -log of the probability given to the label, from the logit (stable)
  const float xl = hit ? -lg.z : lg.z;
  const float bce = active ? fmaxf(xl, 0.0f) + log1pf(expf(-fabsf(xl))) : 0.f;
  const float sq = wave_sum(fmaf(hit_weight, bce, e0 * e0 + e1 * e1));
  const float gb0 = wave_sum(dz40), gb1 = wave_sum(dz41), gb2 = wave_sum(dz42);
  if (lane == 0) {
    loss_part[blockIdx.x] = sq;
    P[OFF_B4] = gb0;
    P[OFF_B4 + 1] = gb1;
    P[OFF_B4 + 2] = gb2;
  }
  // W4 gradient, and dz3 = (W4^T dz4) * gelu'(z3)
#pragma unroll 1
  for (int j = 0; j < NH; j++) {
    const float z = s_z[(2 * NH + j) * 64 + lane];
    const float h = bf16r(gelu(z));
    const float g0 = wave_sum(dz40 * h), g1 = wave_sum(dz41 * h), g2 = wave_sum(dz42 * h);
    const float dh = fmaf(bf16r(w[OFF_W4 + j]), dz40,
                          fmaf(bf16r(w[OFF_W4 + NH + j]), dz41, bf16r(w[OFF_W4 + 2 * NH + j]) * dz42));
    const float dzv = dh * gelu_grad(z);
    s_dzu[j * 64 + lane] = (__bf16)dzv;
    s_dzp[lane * DZP + j] = (__bf16)dzv;
    const float gb = wave_sum(dzv);
    if (lane == 0) {
      P[OFF_W4 + j] = g0;
      P[OFF_W4 + NH + j] = g1;
      P[OFF_W4 + 2 * NH + j] = g2;
      P[OFF_B3 + j] = gb;
    }
  }
  wave_lds_sync();
  // z3 is consumed: its slice holds dh_prev from here on (69.5 KB of LDS per
  // block, so two blocks share a compute unit)
  float* const s_dh = s_z + 2 * NH * 64;
  // layers 3 and 2: dW = dz h_prev^T (k = pairs), dh_prev = W^T dz, dz_prev = dh_prev * gelu'(z_prev)
#pragma unroll 1
  for (int layer = 2; layer >= 1; layer--) {
    const int offw = layer == 2 ? OFF_W3 : OFF_W2;
    const int tt = layer == 2 ? T_T3 : T_T2;
    const float* zprev = s_z + (layer - 1) * NH * 64;
#pragma unroll 1
    for (int ob = 0; ob < 4; ob++) {
#pragma unroll 1
      for (int nb = 0; nb < 4; nb++) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kb = 0; kb < 2; kb++) {
          const bf16x8 a = *reinterpret_cast<const bf16x8*>(s_dzu + (16 * ob + m) * 64 + 32 * kb + 8 * g);
          bf16x8 b;
#pragma unroll
          for (int i = 0; i < 8; i++) b[i] = (__bf16)gelu(zprev[(16 * nb + m) * 64 + 32 * kb + 8 * g + i]);
          acc = mfma(a, b, acc);
        }
#pragma unroll
        for (int i = 0; i < 4; i++) P[offw + (16 * ob + 4 * g + i) * NH + 16 * nb + m] = acc[i];
      }
    }
#pragma unroll 1
    for (int ob = 0; ob < 4; ob++) {
#pragma unroll 1
      for (int nb = 0; nb < 4; nb++) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kb = 0; kb < 2; kb++) {
          const bf16x8 b = *reinterpret_cast<const bf16x8*>(s_dzp + (16 * nb + m) * DZP + 32 * kb + 8 * g);
          acc = mfma(tiles[(tt + ob * 2 + kb) * 64 + lane], b, acc);
        }
#pragma unroll
        for (int i = 0; i < 4; i++) s_dh[(16 * ob + 4 * g + i) * 64 + 16 * nb + m] = acc[i];
      }
    }
    wave_lds_sync();
    const int offb = layer == 2 ? OFF_B2 : OFF_B1;
#pragma unroll 1
    for (int j = 0; j < NH; j++) {
      const float dzv = s_dh[j * 64 + lane] * gelu_grad(zprev[j * 64 + lane]);
      s_dzu[j * 64 + lane] = (__bf16)dzv;
      s_dzp[lane * DZP + j] = (__bf16)dzv;
      const float gb = wave_sum(dzv);
      if (lane == 0) P[offb + j] = gb;
    }
    wave_lds_sync();
  }
  // layer 1: dW1 = dz1 x^T (k = pairs), features 0..19
  const __bf16* xs = reinterpret_cast<const __bf16*>(s_x);
#pragma unroll 1
  for (int ob = 0; ob < 4; ob++) {
#pragma unroll 1
    for (int nb = 0; nb < 2; nb++) {
      f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int kb = 0; kb < 2; kb++) {
        const bf16x8 a = *reinterpret_cast<const bf16x8*>(s_dzu + (16 * ob + m) * 64 + 32 * kb + 8 * g);
        bf16x8 b;
#pragma unroll
        for (int i = 0; i < 8; i++) b[i] = xs[(32 * kb + 8 * g + i) * 32 + 16 * nb + m];
        acc = mfma(a, b, acc);
      }
      const int fcol = 16 * nb + m;
      if (fcol < NIN) {
#pragma unroll
        for (int i = 0; i < 4; i++) P[OFF_W1 + (16 * ob + 4 * g + i) * NIN + fcol] = acc[i];
      }
    }
  }
}

// Gradient reduction in two fixed-order stages (deterministic): group g of
// kReduceGroup blocks sums its partials into stage[g] (NPARAM values, then the
// loss parts' sum), then k_reduce sums the groups into grad and adds the loss.
constexpr int kReduceGroup = 16;

__global__ void k_reduce_groups(const float* __restrict__ partial, int nblocks, const float* __restrict__ loss_part,
                                float* __restrict__ stage) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x, g = blockIdx.y;
  const int b0 = g * kReduceGroup, b1 = min(nblocks, b0 + kReduceGroup);
  if (p < NPARAM) {
    float s = 0.f;
    for (int b = b0; b < b1; b++) s += partial[(int64_t)b * NPARAM + p];
    stage[(int64_t)g * (NPARAM + 1) + p] = s;
  } else if (p == NPARAM) {
    float s = 0.f;
    for (int b = b0; b < b1; b++) s += loss_part[b];
    stage[(int64_t)g * (NPARAM + 1) + NPARAM] = s;
  }
}

__global__ void k_reduce(const float* __restrict__ stage, int ngroups, float* __restrict__ grad,
                         float* __restrict__ loss) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p > NPARAM) return;
  float s = 0.f;
  for (int g = 0; g < ngroups; g++) s += stage[(int64_t)g * (NPARAM + 1) + p];
  if (p < NPARAM) grad[p] = s;
  else loss[0] += s;
}

__global__ void k_adam(float* __restrict__ w, float* __restrict__ m, float* __restrict__ v,
                       const float* __restrict__ g, float lr, float b1, float b2, float eps, float bc1, float bc2) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= NPARAM) return;
  const float gg = g[p];
  const float mm = b1 * m[p] + (1.0f - b1) * gg;
  const float vv = b2 * v[p] + (1.0f - b2) * gg * gg;
  m[p] = mm;
  v[p] = vv;
  w[p] -= lr * (mm / bc1) / (sqrtf(vv / bc2) + eps);
}

// Epoch shuffle on the device: perm = a keyed pseudo-random permutation of
// [0, n).  A four-round balanced Feistel network permutes [0, 4^h) with
// 4^h >= n; cycle-walking (apply it again while the value is >= n) restricts
// it to [0, n).  The walk ends: the cycle through i < n returns to i.
__device__ __forceinline__ uint32_t feistel_round(uint32_t x, uint32_t k) {
  uint32_t z = (x ^ k) * 0x9E3779B1u;
  z ^= z >> 15;
  z *= 0x85EBCA77u;
  z ^= z >> 13;
  return z;
}

__device__ __forceinline__ uint32_t feistel(uint32_t x, int h, const uint32_t (&key)[4]) {
  const uint32_t mask = (1u << h) - 1u;
  uint32_t l = x >> h, r = x & mask;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const uint32_t t = l ^ (feistel_round(r, key[k]) & mask);
    l = r;
    r = t;
  }
  return l << h | r;
}

__global__ void k_permutation(int* __restrict__ perm, int64_t n, int h, uint32_t k0, uint32_t k1, uint32_t k2,
                              uint32_t k3) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t key[4] = {k0, k1, k2, k3};
  uint32_t y = feistel((uint32_t)i, h, key);
  while (y >= (uint64_t)n) y = feistel(y, h, key);
  perm[i] = (int)y;
}

int pack(rtn_mlp* net, hipStream_t st) {
  hipLaunchKernelGGL(k_pack, dim3((NTILE * 64 + 255) / 256), dim3(256), 0, st, net->w.ptr, net->tiles.ptr,
                     net->bias.ptr);
  RTN_HIP(hipGetLastError());
  return RTN_OK;
}

int64_t reduce_groups(int64_t blocks) { return (blocks + kReduceGroup - 1) / kReduceGroup; }

// partial: blocks x NPARAM gradient partials, the blocks' loss parts, then
// the reduction's group sums (groups x (NPARAM + 1)).
int ensure_partial(rtn_mlp* net, int64_t blocks) {
  if (blocks <= net->partial_blocks) return RTN_OK;
  RTN_TRY(net->partial.alloc((size_t)blocks * NPARAM + blocks + (size_t)reduce_groups(blocks) * (NPARAM + 1)));
  net->partial_blocks = blocks;
  return RTN_OK;
}

// Gradient of the batch [base, end) of the (optionally permuted) pairs into net->grad.
int batch_gradient(rtn_mlp* net, const float* feats, const float* targets, const int* perm, int64_t base, int64_t end,
                   float hit_weight, hipStream_t st) {
  const int64_t nb = end - base;
  const int blocks = (int)((nb + 63) / 64);
  RTN_TRY(ensure_partial(net, blocks));
  float* loss_part = net->partial.ptr + (size_t)net->partial_blocks * NPARAM;
  hipLaunchKernelGGL(k_train, dim3(blocks), dim3(64), 0, st, net->tiles.ptr, net->bias.ptr, net->w.ptr, feats, targets,
                     perm, base, end, 1.0f / (float)nb, hit_weight, net->partial.ptr, loss_part);
  RTN_HIP(hipGetLastError());
  float* stage = loss_part + net->partial_blocks;
  const int groups = (int)reduce_groups(blocks);
  hipLaunchKernelGGL(k_reduce_groups, dim3((NPARAM + 1 + 255) / 256, groups), dim3(256), 0, st, net->partial.ptr,
                     blocks, loss_part, stage);
  hipLaunchKernelGGL(k_reduce, dim3((NPARAM + 1 + 255) / 256), dim3(256), 0, st, stage, groups, net->grad.ptr,
                     net->loss.ptr);
  RTN_HIP(hipGetLastError());
  return RTN_OK;
}

uint64_t splitmix(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

}  // namespace

namespace rtn {
const void* mlp_tiles(const rtn_mlp* net) { return net->tiles.ptr; }
int mlp_kind(const rtn_mlp* net) { return net->kind; }
const float* mlp_bias(const rtn_mlp* net) { return net->bias.ptr; }
}  // namespace rtn

extern "C" int rtn_mlp_create(rtn_mlp** out, uint64_t seed, int kind) {
  RTN_CHECK(out, RTN_EINVAL, "rtn_mlp_create: null output");
  *out = nullptr;
  RTN_CHECK(kind >= 0 && kind < FEAT_KINDS, RTN_EINVAL, "rtn_mlp_create: unknown feature kind %d", kind);
  rtn_mlp* net = new rtn_mlp();
  net->kind = kind;
  int rc = RTN_OK;
  if ((rc = net->w.alloc(NPARAM)) || (rc = net->m.alloc(NPARAM)) || (rc = net->v.alloc(NPARAM)) ||
      (rc = net->grad.alloc(NPARAM)) || (rc = net->tiles.alloc((size_t)NTILE * 64)) ||
      (rc = net->bias.alloc(NBIAS)) || (rc = net->loss.alloc(1))) {
    delete net;
    return rc;
  }
  // Glorot-uniform weights, zero biases
  std::vector<float> h(NPARAM, 0.0f);
  uint64_t s = seed;
  auto fill = [&](int off, int fo, int fi) {
    const double a = std::sqrt(6.0 / (fo + fi));
    for (int i = 0; i < fo * fi; i++) h[off + i] = (float)(a * (2.0 * ((splitmix(s) >> 11) * 0x1.0p-53) - 1.0));
  };
  fill(OFF_W1, NH, NIN);
  fill(OFF_W2, NH, NH);
  fill(OFF_W3, NH, NH);
  fill(OFF_W4, NOUT, NH);
  *out = net;
  return rtn_mlp_set_params(net, h.data());
}

extern "C" void rtn_mlp_destroy(rtn_mlp* net) { delete net; }

extern "C" int rtn_mlp_kind(const rtn_mlp* net) { return net ? net->kind : -1; }

extern "C" int rtn_mlp_set_params(rtn_mlp* net, const float* params) {
  RTN_CHECK(net && params, RTN_EINVAL, "rtn_mlp_set_params: null argument");
  for (int i = 0; i < NPARAM; i++) RTN_CHECK(std::isfinite(params[i]), RTN_EINVAL, "rtn_mlp_set_params: param %d not finite", i);
  RTN_TRY(net->w.upload(params, NPARAM));
  RTN_HIP(hipMemset(net->m.ptr, 0, NPARAM * sizeof(float)));
  RTN_HIP(hipMemset(net->v.ptr, 0, NPARAM * sizeof(float)));
  net->step = 0;
  RTN_TRY(pack(net, nullptr));
  RTN_HIP(hipDeviceSynchronize());
  return RTN_OK;
}

extern "C" int rtn_mlp_get_params(const rtn_mlp* net, float* params) {
  RTN_CHECK(net && params, RTN_EINVAL, "rtn_mlp_get_params: null argument");
  RTN_HIP(hipDeviceSynchronize());
  return net->w.download(params, NPARAM);
}

extern "C" int rtn_mlp_pair_features(const rtn_scene* s, const float* const rays[6], int64_t n_rays,
                                     const int* pair_ray, const int* pair_patch, int64_t n, int kind, float* feats,
                                     void* stream) {
  RTN_CHECK(s && s->built, RTN_EINVAL, "rtn_mlp_pair_features: scene not built");
  RTN_CHECK(rays && pair_ray && pair_patch && feats && n >= 0, RTN_EINVAL, "rtn_mlp_pair_features: bad arguments");
  RTN_CHECK(kind >= 0 && kind < FEAT_KINDS, RTN_EINVAL, "rtn_mlp_pair_features: unknown feature kind %d", kind);
  for (int k = 0; k < 6; k++) RTN_CHECK(rays[k], RTN_EINVAL, "rtn_mlp_pair_features: null ray component %d", k);
  if (n == 0) return RTN_OK;
  hipStream_t st = (hipStream_t)stream;
  DevBuf<int> bad;
  RTN_TRY(bad.alloc(1));
  RTN_HIP(hipMemsetAsync(bad.ptr, 0, sizeof(int), st));
  hipLaunchKernelGGL(k_features, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, s->bounds.ptr, s->cp.ptr,
                     s->cp_off.ptr, s->pq.ptr, (int)s->n_patches, rays[0], rays[1], rays[2], rays[3], rays[4], rays[5], n_rays,
                     pair_ray, pair_patch, n, kind, feats, bad.ptr);
  RTN_HIP(hipGetLastError());
  int hb = 0;
  RTN_HIP(hipMemcpyAsync(&hb, bad.ptr, sizeof(int), hipMemcpyDeviceToHost, st));
  RTN_HIP(hipStreamSynchronize(st));
  RTN_CHECK(!hb, RTN_EINVAL, "rtn_mlp_pair_features: pair index out of range");
  return RTN_OK;
}

extern "C" int rtn_mlp_forward(const rtn_mlp* net, const float* feats, const int* valid, int64_t n, float* uv,
                               float* hit, void* stream) {
  RTN_CHECK(net && feats && uv && n >= 0, RTN_EINVAL, "rtn_mlp_forward: bad arguments");
  if (n == 0) return RTN_OK;
  const int64_t per = 64 * FWD_WAVES;
  hipLaunchKernelGGL(k_forward, dim3((unsigned)((n + per - 1) / per)), dim3(per), 0, (hipStream_t)stream,
                     net->tiles.ptr, net->bias.ptr, feats, valid, n, uv, hit);
  RTN_HIP(hipGetLastError());
  return RTN_OK;
}

extern "C" int rtn_mlp_gradient(rtn_mlp* net, const float* feats, const float* targets, int64_t n, float hit_weight,
                                float* grad, float* loss) {
  RTN_CHECK(net && feats && targets && grad && n > 0, RTN_EINVAL, "rtn_mlp_gradient: bad arguments");
  RTN_CHECK(n <= (1 << 22), RTN_EINVAL, "rtn_mlp_gradient: at most 4194304 pairs per call");
  RTN_HIP(hipMemset(net->loss.ptr, 0, sizeof(float)));
  RTN_TRY(batch_gradient(net, feats, targets, nullptr, 0, n, hit_weight, nullptr));
  RTN_HIP(hipDeviceSynchronize());
  RTN_TRY(net->grad.download(grad, NPARAM));
  if (loss) {
    float l = 0.f;
    RTN_TRY(net->loss.download(&l, 1));
    *loss = l / (float)n;
  }
  return RTN_OK;
}

extern "C" void rtn_mlp_train_params_default(rtn_mlp_train_params* tp) {
  if (!tp) return;
  memset(tp, 0, sizeof(*tp));
  tp->epochs = 4;
  tp->batch = 32768;  // 512 single-wave blocks: two per compute unit, one resident round
  tp->lr = 6e-3f;
  tp->lr_final = 1e-4f;
  tp->beta1 = 0.9f;
  tp->beta2 = 0.999f;
  tp->eps = 1e-8f;
  tp->hit_weight = 0.05f;
  tp->seed = 1;
}

extern "C" int rtn_mlp_train(rtn_mlp* net, const float* feats, const float* targets, int64_t n,
                             const rtn_mlp_train_params* tp, float* epoch_loss, void* stream) {
  RTN_CHECK(net && feats && targets && tp && n > 0, RTN_EINVAL, "rtn_mlp_train: bad arguments");
  RTN_CHECK(n < ((int64_t)1 << 31), RTN_EINVAL, "rtn_mlp_train: too many pairs");
  RTN_CHECK(tp->epochs >= 1 && tp->batch >= 64 && tp->batch <= (1 << 20) && tp->lr > 0 && tp->hit_weight >= 0,
            RTN_EINVAL,
            "rtn_mlp_train: bad training parameters");
  hipStream_t st = (hipStream_t)stream;
  DevBuf<int> perm;
  RTN_TRY(perm.alloc((size_t)n));
  int h = 1;  // Feistel half width: 4^h >= n
  while (((int64_t)1 << (2 * h)) < n) h++;
  uint64_t ks = tp->seed;
  const int64_t steps_per_epoch = (n + tp->batch - 1) / tp->batch;
  const int64_t total = steps_per_epoch * tp->epochs;
  int64_t k = 0;
  for (int e = 0; e < tp->epochs; e++) {
    uint32_t key[4];
    for (uint32_t& x : key) x = (uint32_t)splitmix(ks);
    hipLaunchKernelGGL(k_permutation, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, perm.ptr, n, h, key[0],
                       key[1], key[2], key[3]);
    RTN_HIP(hipGetLastError());
    RTN_HIP(hipMemsetAsync(net->loss.ptr, 0, sizeof(float), st));
    for (int64_t b = 0; b < n; b += tp->batch, k++) {
      const int64_t end = std::min<int64_t>(n, b + tp->batch);
      RTN_TRY(batch_gradient(net, feats, targets, perm.ptr, b, end, tp->hit_weight, st));
      // cosine decay from lr to lr_final
      const double c = 0.5 * (1.0 + std::cos(M_PI * (double)k / (double)std::max<int64_t>(1, total - 1)));
      const float lr = (float)(tp->lr_final + (tp->lr - tp->lr_final) * c);
      net->step++;
      const float bc1 = 1.0f - (float)std::pow((double)tp->beta1, (double)net->step);
      const float bc2 = 1.0f - (float)std::pow((double)tp->beta2, (double)net->step);
      hipLaunchKernelGGL(k_adam, dim3((NPARAM + 255) / 256), dim3(256), 0, st, net->w.ptr, net->m.ptr, net->v.ptr,
                         net->grad.ptr, lr, tp->beta1, tp->beta2, tp->eps, bc1, bc2);
      RTN_HIP(hipGetLastError());
      RTN_TRY(pack(net, st));
    }
    float l = 0.f;
    RTN_HIP(hipMemcpyAsync(&l, net->loss.ptr, sizeof(float), hipMemcpyDeviceToHost, st));
    RTN_HIP(hipStreamSynchronize(st));
    if (epoch_loss) epoch_loss[e] = l / (float)n;
  }
  return RTN_OK;
}
