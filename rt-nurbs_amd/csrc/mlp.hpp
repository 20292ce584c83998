// Learned Newton seed: a 20 -> 64 -> 64 -> 64 -> 3 perceptron (GELU hidden
// units: tanh form in training, its polynomial fit gelu_infer in inference;
// sigmoid outputs) mapping a (ray, patch) pair to a starting (u,v) and
// the probability that Newton converges on the pair (a hit).
//
// Inference runs per wave on the matrix cores: v_mfma_f32_16x16x32_bf16 with
// BF16 weights and activations and FP32 accumulation.  The pairs of a wave are
// the N dimension (16 per block, up to 4 blocks), hidden units the M
// dimension, so one layer is D^T = W X^T + b.  The output fragment of a layer
// (lane l holds units 4*(l/16)+i, i<4, of pair l%16) becomes the next layer's
// B fragment (lane l holds k = 8*(l/16)+i of pair l%16) without any data
// movement: slot i of lane group g carries unit perm_unit(g,i) of each 32-unit
// block, and the weight tiles store their k index in that same order.
//
// Operand layouts of v_mfma_f32_16x16x32_bf16 (wave64):
//   A (16 x 32): lane l holds A[l%16][8*(l/16) + i], i = 0..7
//   B (32 x 16): lane l holds B[8*(l/16) + i][l%16]
//   D (16 x 16): lane l holds D[4*(l/16) + i][l%16], i = 0..3
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "rtn_mem.hpp"

namespace rtn {
namespace mlp {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// Inputs: the 20 pair features of one of the feature kinds below.  Outputs:
// logits of u, v and of the probability that Newton converges (a hit).
constexpr int NIN = 20, NH = 64, NOUT = 3;
enum FeatureKind : int {
  FEAT_PROJECTED = 0,  // the control net projected on the ray planes (pair_features)
  FEAT_BOX = 1,        // ray, corners and degrees in the patch box's frame (box_features)
  FEAT_KINDS = 2
};

// FP32 master parameters, row-major (out x in) weights then bias per layer.
constexpr int OFF_W1 = 0;
constexpr int OFF_B1 = OFF_W1 + NH * NIN;
constexpr int OFF_W2 = OFF_B1 + NH;
constexpr int OFF_B2 = OFF_W2 + NH * NH;
constexpr int OFF_W3 = OFF_B2 + NH;
constexpr int OFF_B3 = OFF_W3 + NH * NH;
constexpr int OFF_W4 = OFF_B3 + NH;
constexpr int OFF_B4 = OFF_W4 + NOUT * NH;
constexpr int NPARAM = OFF_B4 + NOUT;  // 9794

// BF16 A-operand tiles, 64 lanes x 8 values each.  Forward: layer 1 (4 row
// blocks), layers 2 and 3 (4 row blocks x 2 k blocks, tile ob*2+kb), output
// layer (rows 0,1 real, 2 k blocks).  Backward: W2^T and W3^T (natural k order).
constexpr int T_L1 = 0, T_L2 = 4, T_L3 = 12, T_L4 = 20, T_T2 = 22, T_T3 = 30, NTILE = 38;
constexpr int NTILE_FWD = 22;
// FP32 biases as accumulator seeds: b1 | b2 | b3 | b4 padded to 16 rows
// (rows 0..2 real).
constexpr int BB1 = 0, BB2 = 64, BB3 = 128, BB4 = 192, NBIAS = 208;

// Unit carried by slot i (0..7) of lane group g within a 32-unit k block.
__host__ __device__ constexpr int perm_unit(int g, int i) { return i < 4 ? 4 * g + i : 16 + 4 * g + (i - 4); }

// GELU, tanh form: 0.5 x (1 + tanh(k0 (x + k1 x^3))) = x * sigmoid(2 k0 (x + k1 x^3)).
constexpr float GK0 = 0.7978845608028654f, GK1 = 0.044715f;
constexpr float LOG2E = 1.4426950408889634f;

__device__ __forceinline__ float fast_sigmoid(float z) {
  return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(-z * LOG2E));
}
__device__ __forceinline__ float gelu(float x) {
  constexpr float C = -2.0f * GK0 * LOG2E;  // sigmoid(z) = 1 / (1 + 2^(-z log2 e))
  const float e = __builtin_amdgcn_exp2f(x * C * fmaf(GK1, x * x, 1.0f));
  return x * __builtin_amdgcn_rcpf(1.0f + e);
}
// Inference GELU: x * clamp(0.5 + x Q(x^2), 0, 1) with Q of degree 4, the
// minimax fit of gelu(x)/x - 0.5 on [0, 3.4] held to 3.4 Q(3.4^2) = 0.5, so the
// clamp takes over at |x| = 3.4 (x above, 0 below).  Within 1.075e-3 of the
// tanh form on the whole line (the BF16 rounding that follows moves an
// activation in [0.5, 1) by up to 1.95e-3): full-rate FMAs in place of an
// exponential and a reciprocal.  The training passes keep the tanh form (its
// gradient is gelu_grad).
__device__ __forceinline__ float gelu_infer(float x) {
  const float x2 = x * x;
  float t = fmaf(1.0574128852131707e-05f, x2, -4.152066351606772e-04f);
  t = fmaf(t, x2, 6.703262281516497e-03f);
  t = fmaf(t, x2, -5.970189424328037e-02f);
  t = fmaf(t, x2, 3.940125455699649e-01f);
  return x * __builtin_amdgcn_fmed3f(fmaf(x, t, 0.5f), 0.0f, 1.0f);
}
// d gelu / dx = s + 2 x s (1 - s) k0 (1 + 3 k1 x^2), s = sigmoid(2 k0 (x + k1 x^3))
__device__ __forceinline__ float gelu_grad(float x) {
  const float x2 = x * x;
  const float s = fast_sigmoid(2.0f * GK0 * x * fmaf(GK1, x2, 1.0f));
  return fmaf(2.0f * x * s * (1.0f - s), GK0 * fmaf(3.0f * GK1, x2, 1.0f), s);
}
__device__ __forceinline__ float bf16r(float x) { return (float)(__bf16)x; }

// Input features of kind FEAT_PROJECTED (20 values).  On the two ray planes the ray is
// the origin and the patch a 2D map of (u,v) whose control points are the
// projected control points: the seed is where that map reaches the origin.
// The features sample the projected net on a 3 x 3 grid of control indices,
// rows {0, P/2, P} by columns {0, Q/2, Q} (an odd degree's middle averages its
// two central indices), each sample the mean of the dehomogenised points at
// those indices, projected as (n1.(x - o), n2.(x - o)) and divided by the
// patch box's largest half extent hs:
//   f[2 (3a + b)], f[2 (3a + b) + 1]  sample (row a, column b), a, b = 0..2
//   f[18], f[19]                      degrees, (p - 3.5) / 1.5, (q - 3.5) / 1.5
// n1, n2: unit ray-plane normals (perpendicular to the ray direction).
// Reciprocals are hardware approximations: the features end in BF16.
template <class NP = const float4*, class BP = const float*>
__device__ __forceinline__ void pair_features(float ox, float oy, float oz, float n1x, float n1y, float n1z,
                                              float n2x, float n2y, float n2z, NP net, int P, int Q, BP bnd,
                                              float* f) {
  const float hs = fmaxf(fmaxf(ld1(bnd, 3) - ld1(bnd, 0), ld1(bnd, 4) - ld1(bnd, 1)),
                         fmaxf(ld1(bnd, 5) - ld1(bnd, 2), 2e-20f)) * 0.5f;
  const float is = __builtin_amdgcn_rcpf(hs);
  const int ri[4] = {0, P >> 1, (P + 1) >> 1, P}, ci[4] = {0, Q >> 1, (Q + 1) >> 1, Q};
  float a1[4][4], a2[4][4];  // projected points at rows ri x columns ci
#pragma unroll
  for (int r = 0; r < 4; r++)
#pragma unroll
    for (int c = 0; c < 4; c++) {
      const float4 x = ld4(net, ri[r] * (Q + 1) + ci[c]);
      const float iw = __builtin_amdgcn_rcpf(x.w);
      const float ex = fmaf(x.x, iw, -ox), ey = fmaf(x.y, iw, -oy), ez = fmaf(x.z, iw, -oz);
      a1[r][c] = fmaf(n1x, ex, fmaf(n1y, ey, n1z * ez)) * is;
      a2[r][c] = fmaf(n2x, ex, fmaf(n2y, ey, n2z * ez)) * is;
    }
  // sample rows/columns as index pairs into ri / ci: {0,0}, {1,2}, {3,3}
  const int lo[3] = {0, 1, 3}, hi[3] = {0, 2, 3};
#pragma unroll
  for (int sa = 0; sa < 3; sa++)
#pragma unroll
    for (int sb = 0; sb < 3; sb++) {
      const int k = 2 * (3 * sa + sb);
      f[k] = 0.25f * ((a1[lo[sa]][lo[sb]] + a1[lo[sa]][hi[sb]]) + (a1[hi[sa]][lo[sb]] + a1[hi[sa]][hi[sb]]));
      f[k + 1] = 0.25f * ((a2[lo[sa]][lo[sb]] + a2[lo[sa]][hi[sb]]) + (a2[hi[sa]][lo[sb]] + a2[hi[sa]][hi[sb]]));
    }
  f[18] = ((float)P - 3.5f) * (1.0f / 1.5f);
  f[19] = ((float)Q - 3.5f) * (1.0f / 1.5f);
}

// Input features of kind FEAT_BOX: the ray origin / direction in the local
// frame of the patch's box (centre c, largest half extent hs), the degrees and
// the four corner control points.  The origin is moved along the ray to
// where it enters the box (t_e = the slab entry, at least 0), so it lies in
// the box's scale:
//   f[0..2]   (o + t_e d - c) / hs          f[3..5]  d / |d|
//   f[6..17]  (corner - c) / hs for the dehomogenised corners P00, Pp0, P0q, Ppq
//   f[18], f[19]  degrees, (p - 3.5) / 1.5, (q - 3.5) / 1.5
// Slab terms with a zero direction component are NaN or infinite; fminf /
// fmaxf drop NaN operands, an infinite entry (a ray parallel to a slab it
// lies outside) counts as 0, and the entry-point features are clamped to
// [-4, 4].  Reciprocals are hardware approximations.
template <class NP = const float4*, class BP = const float*>
__device__ __forceinline__ void box_features(float ox, float oy, float oz, float dx, float dy, float dz, NP net,
                                             int P, int Q, BP bnd, float* f) {
  const float lo[3] = {ld1(bnd, 0), ld1(bnd, 1), ld1(bnd, 2)}, hi[3] = {ld1(bnd, 3), ld1(bnd, 4), ld1(bnd, 5)};
  const float o[3] = {ox, oy, oz}, d[3] = {dx, dy, dz};
  float te = 0.0f, hs = 2e-20f, c[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const float id = __builtin_amdgcn_rcpf(d[k]);
    const float a = (lo[k] - o[k]) * id, b = (hi[k] - o[k]) * id;
    te = fmaxf(te, fminf(a, b));
    c[k] = 0.5f * (lo[k] + hi[k]);
    hs = fmaxf(hs, 0.5f * (hi[k] - lo[k]));
  }
  if (!(te < 3.0e38f)) te = 0.0f;
  const float is = __builtin_amdgcn_rcpf(hs);
  const float dn = __builtin_amdgcn_rsqf(fmaf(dx, dx, fmaf(dy, dy, dz * dz)));
#pragma unroll
  for (int k = 0; k < 3; k++) {
    f[k] = fminf(fmaxf((fmaf(te, d[k], o[k]) - c[k]) * is, -4.0f), 4.0f);
    f[3 + k] = d[k] * dn;
  }
  const int at[4] = {0, P * (Q + 1), Q, P * (Q + 1) + Q};  // P00, Pp0, P0q, Ppq
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const float4 x = ld4(net, at[k]);
    const float iw = __builtin_amdgcn_rcpf(x.w);
    f[6 + 3 * k] = (x.x * iw - c[0]) * is;
    f[7 + 3 * k] = (x.y * iw - c[1]) * is;
    f[8 + 3 * k] = (x.z * iw - c[2]) * is;
  }
  f[18] = ((float)P - 3.5f) * (1.0f / 1.5f);
  f[19] = ((float)Q - 3.5f) * (1.0f / 1.5f);
}

// The NIN inputs of a pair for a feature kind.
template <class NP = const float4*, class BP = const float*>
__device__ __forceinline__ void kind_features(int kind, float ox, float oy, float oz, float dx, float dy, float dz,
                                              float n1x, float n1y, float n1z, float n2x, float n2y, float n2z,
                                              NP net, int P, int Q, BP bnd, float f[NIN]) {
  if (kind == FEAT_BOX) box_features(ox, oy, oz, dx, dy, dz, net, P, Q, bnd, f);
  else pair_features(ox, oy, oz, n1x, n1y, n1z, n2x, n2y, n2z, net, P, Q, bnd, f);
}

__device__ __forceinline__ f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ f32x4 bias4(const float* __restrict__ bias, int at) {
  const float4 b = *reinterpret_cast<const float4*>(bias + at);
  return f32x4{b.x, b.y, b.z, b.w};
}

// Orders this wave's LDS writes before its later LDS reads by other lanes.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Forward pass for nblk blocks of 16 pairs.  s_x[r][c] holds the features of
// pair row r as BF16 (c = 8-value column group, inputs 20..31 zero).  The
// output logits of pair r go to s_logit[r] (x, y, z: u, v, hit).  KEEP (training) also stores the
// FP32 pre-activations of the three hidden layers, s_z[layer][unit][pair], and
// uses the tanh GELU of training; inference uses gelu_infer.
// Must be called by all 64 lanes with nblk wave-uniform.
template <bool KEEP>
__device__ __forceinline__ void forward_wave(const bf16x8* __restrict__ tiles, const float* __restrict__ bias,
                                             const bf16x8 (*s_x)[4], int nblk, int lane, float4* s_logit,
                                             float* s_z) {
  const int m = lane & 15, g = lane >> 4;
#pragma unroll 1
  for (int nb = 0; nb < nblk; nb++) {
    f32x4 d[4];
    bf16x8 h[2];
    const bf16x8 x = s_x[16 * nb + m][g];
    // keep each layer's tile loads next to its products (register pressure): an
    // opaque offset stops the loads from being hoisted, and offsetting (rather
    // than hiding the pointer) keeps them global-memory loads
    int toff = 0;
    asm volatile("" : "+v"(toff));
    const bf16x8* tp = tiles + toff;
#pragma unroll
    for (int ob = 0; ob < 4; ob++) d[ob] = mfma(tp[(T_L1 + ob) * 64 + lane], x, bias4(bias, BB1 + 16 * ob + 4 * g));
#pragma unroll
    for (int layer = 0; layer < 3; layer++) {
      if (KEEP) {
#pragma unroll
        for (int ob = 0; ob < 4; ob++)
#pragma unroll
          for (int i = 0; i < 4; i++) s_z[(layer * NH + 16 * ob + 4 * g + i) * 64 + 16 * nb + m] = d[ob][i];
      }
#pragma unroll
      for (int kb = 0; kb < 2; kb++)
#pragma unroll
        for (int i = 0; i < 8; i++) {
          const float z = d[2 * kb + (i >> 2)][i & 3];
          h[kb][i] = (__bf16)(KEEP ? gelu(z) : gelu_infer(z));
        }
      if (layer == 2) break;
      const int tl = layer == 0 ? T_L2 : T_L3;
      const int bb = layer == 0 ? BB2 : BB3;
      asm volatile("" : "+v"(toff));
      tp = tiles + toff;
#pragma unroll
      for (int ob = 0; ob < 4; ob++) {
        f32x4 acc = bias4(bias, bb + 16 * ob + 4 * g);
        acc = mfma(tp[(tl + ob * 2) * 64 + lane], h[0], acc);
        d[ob] = mfma(tp[(tl + ob * 2 + 1) * 64 + lane], h[1], acc);
      }
    }
    asm volatile("" : "+v"(toff));
    tp = tiles + toff;
    f32x4 o = bias4(bias, BB4 + 4 * g);
    o = mfma(tp[(T_L4 + 0) * 64 + lane], h[0], o);
    o = mfma(tp[(T_L4 + 1) * 64 + lane], h[1], o);
    if (g == 0) s_logit[16 * nb + m] = make_float4(o[0], o[1], o[2], 0.0f);
  }
}

// Inference pass of the seed network for nblk blocks of 16 pair rows (rows of
// three 8-value groups), each lane receiving the logits of pair row `row`.
// The weight loads are issued ahead of their products: the first layer's
// tiles and biases once for all blocks, every later layer's before the
// previous layer's GELU, so their latency overlaps that work.  Must be called
// by all 64 lanes with nblk wave-uniform.
__device__ __forceinline__ void forward_seed(const bf16x8* __restrict__ tiles, const float* __restrict__ bias,
                                             const bf16x8 (*s_x)[3], int nblk, int lane, int row, float& lu,
                                             float& lv, float& lh) {
  const int m = lane & 15, g = lane >> 4;
  bf16x8 w1[4];
  f32x4 b1[4];
#pragma unroll
  for (int ob = 0; ob < 4; ob++) {
    w1[ob] = tiles[(T_L1 + ob) * 64 + lane];
    b1[ob] = bias4(bias, BB1 + 16 * ob + 4 * g);
  }
#pragma unroll 1
  for (int nb = 0; nb < nblk; nb++) {
    int toff = 0;  // opaque offset: later layers' tiles are loaded per block, not held across blocks
    asm volatile("" : "+v"(toff));
    const bf16x8* tp = tiles + toff;
    bf16x8 x = {0, 0, 0, 0, 0, 0, 0, 0};
    if (g < 3) x = s_x[16 * nb + m][g];
    f32x4 d[4];
#pragma unroll
    for (int ob = 0; ob < 4; ob++) d[ob] = mfma(w1[ob], x, b1[ob]);
    bf16x8 h[2];
#pragma unroll
    for (int layer = 0; layer < 2; layer++) {
      const int tl = layer == 0 ? T_L2 : T_L3;
      const int bb = layer == 0 ? BB2 : BB3;
      bf16x8 w[8];
      f32x4 b[4];
#pragma unroll
      for (int j = 0; j < 8; j++) w[j] = tp[(tl + j) * 64 + lane];
#pragma unroll
      for (int ob = 0; ob < 4; ob++) b[ob] = bias4(bias, bb + 16 * ob + 4 * g);
#pragma unroll
      for (int kb = 0; kb < 2; kb++)
#pragma unroll
        for (int i = 0; i < 8; i++) h[kb][i] = (__bf16)gelu_infer(d[2 * kb + (i >> 2)][i & 3]);
#pragma unroll
      for (int ob = 0; ob < 4; ob++) d[ob] = mfma(w[2 * ob + 1], h[1], mfma(w[2 * ob], h[0], b[ob]));
    }
    const bf16x8 w40 = tp[(T_L4 + 0) * 64 + lane], w41 = tp[(T_L4 + 1) * 64 + lane];
    const f32x4 b4 = bias4(bias, BB4 + 4 * g);
#pragma unroll
    for (int kb = 0; kb < 2; kb++)
#pragma unroll
      for (int i = 0; i < 8; i++) h[kb][i] = (__bf16)gelu_infer(d[2 * kb + (i >> 2)][i & 3]);
    const f32x4 o = mfma(w41, h[1], mfma(w40, h[0], b4));
    // lane m (< 16) holds the logits of pair 16 nb + m
    const float a = __shfl(o[0], row & 15), bv = __shfl(o[1], row & 15), c = __shfl(o[2], row & 15);
    if ((row >> 4) == nb) {
      lu = a;
      lv = bv;
      lh = c;
    }
  }
}

// forward_seed with each layer's weight tiles loaded next to its products
// instead of ahead: the same products in the same order (bit-identical
// logits) holding fewer registers over the call.  The packet kernel keeps its
// whole ray state live across the call; with this form it spills 44-52 bytes
// per lane instead of 108 and runs 3-4% faster, while the per-lane kernel is
// 3% faster with the early loads of forward_seed.
__device__ __forceinline__ void forward_seed_lean(const bf16x8* __restrict__ tiles,
                                                  const float* __restrict__ bias,
                                                  const bf16x8 (*s_x)[3], int nblk, int lane, int row, float& lu,
                                                  float& lv, float& lh) {
  const int m = lane & 15, g = lane >> 4;
#pragma unroll 1
  for (int nb = 0; nb < nblk; nb++) {
    int toff = 0;  // opaque offset: each layer's tiles are loaded next to its products
    asm volatile("" : "+v"(toff));
    const bf16x8* tp = tiles + toff;
    bf16x8 x = {0, 0, 0, 0, 0, 0, 0, 0};
    if (g < 3) x = s_x[16 * nb + m][g];
    f32x4 d[4];
#pragma unroll
    for (int ob = 0; ob < 4; ob++) d[ob] = mfma(tp[(T_L1 + ob) * 64 + lane], x, bias4(bias, BB1 + 16 * ob + 4 * g));
    bf16x8 h[2];
#pragma unroll
    for (int layer = 0; layer < 2; layer++) {
      const int tl = layer == 0 ? T_L2 : T_L3;
      const int bb = layer == 0 ? BB2 : BB3;
#pragma unroll
      for (int kb = 0; kb < 2; kb++)
#pragma unroll
        for (int i = 0; i < 8; i++) h[kb][i] = (__bf16)gelu_infer(d[2 * kb + (i >> 2)][i & 3]);
      asm volatile("" : "+v"(toff));
      tp = tiles + toff;
#pragma unroll
      for (int ob = 0; ob < 4; ob++)
        d[ob] = mfma(tp[(tl + 2 * ob + 1) * 64 + lane], h[1],
                     mfma(tp[(tl + 2 * ob) * 64 + lane], h[0], bias4(bias, bb + 16 * ob + 4 * g)));
    }
#pragma unroll
    for (int kb = 0; kb < 2; kb++)
#pragma unroll
      for (int i = 0; i < 8; i++) h[kb][i] = (__bf16)gelu_infer(d[2 * kb + (i >> 2)][i & 3]);
    asm volatile("" : "+v"(toff));
    tp = tiles + toff;
    const f32x4 o =
        mfma(tp[(T_L4 + 1) * 64 + lane], h[1], mfma(tp[(T_L4 + 0) * 64 + lane], h[0], bias4(bias, BB4 + 4 * g)));
    const float a = __shfl(o[0], row & 15), bv = __shfl(o[1], row & 15), c = __shfl(o[2], row & 15);
    if ((row >> 4) == nb) {
      lu = a;
      lv = bv;
      lh = c;
    }
  }
}

// Packs a pair's inputs as one BF16 row of s_x (32 values, 20..31 zero).
__device__ __forceinline__ void store_row(bf16x8 (*s_x)[4], int row, const float f[NIN], bool zero) {
  bf16x8 r[4];
#pragma unroll
  for (int c = 0; c < 4; c++)
#pragma unroll
    for (int i = 0; i < 8; i++) {
      const int k = 8 * c + i;
      r[c][i] = (__bf16)((k < NIN && !zero) ? f[k < NIN ? k : 0] : 0.0f);
    }
#pragma unroll
  for (int c = 0; c < 4; c++) s_x[row][c] = r[c];
}

// Input row of one pair (kind_features) written into s_x[row] as three
// 8-value BF16 groups (values 20..23 zero).
template <class NP = const float4*, class BP = const float*>
__device__ __forceinline__ void store_features(bf16x8 (*s_x)[3], int row, int kind, float ox, float oy, float oz,
                                               float dx, float dy, float dz, float n1x, float n1y, float n1z,
                                               float n2x, float n2y, float n2z, NP net, int P, int Q, BP bnd) {
  float f[NIN];
  kind_features(kind, ox, oy, oz, dx, dy, dz, n1x, n1y, n1z, n2x, n2y, n2z, net, P, Q, bnd, f);
#pragma unroll
  for (int c = 0; c < 3; c++) {
    bf16x8 r;
#pragma unroll
    for (int i = 0; i < 8; i++) r[i] = (__bf16)(8 * c + i < NIN ? f[8 * c + i < NIN ? 8 * c + i : 0] : 0.0f);
    s_x[row][c] = r;
  }
}

// Seeds for the active lanes of a wave from their input rows: the active
// pairs' rows were written compacted (rows 0..k-1, see seed_row), so a wave
// with k active lanes runs ceil(k/16) pair blocks.  Called by all 64 lanes;
// u, v and the hit probability ph are written for active lanes only.  LEAN
// selects forward_seed_lean.
template <bool LEAN>
__device__ __forceinline__ void seed_wave_rows(const bf16x8* __restrict__ tiles, const float* __restrict__ bias,
                                               bool active, int row, int n_act, float& u, float& v, float& ph,
                                               const bf16x8 (*s_x)[3], int lane) {
  wave_lds_sync();
  float lu = 0.0f, lv = 0.0f, lh = 0.0f;
  if (LEAN) forward_seed_lean(tiles, bias, s_x, (n_act + 15) >> 4, lane, row, lu, lv, lh);
  else forward_seed(tiles, bias, s_x, (n_act + 15) >> 4, lane, row, lu, lv, lh);
  if (active) {
    u = fast_sigmoid(lu);
    v = fast_sigmoid(lv);
    ph = fast_sigmoid(lh);
  }
  wave_lds_sync();  // the next call's row writes come after this call's reads
}

// Row of this lane in the compacted layout: active lanes first, in lane
// order, then the inactive lanes; n_act = number of active lanes.
__device__ __forceinline__ int seed_row(bool active, int lane, int& n_act) {
  const unsigned long long bal = __ballot(active);
  n_act = __popcll(bal);
  const unsigned long long below = lane ? (~0ull >> (64 - lane)) : 0ull;
  return active ? __popcll(bal & below) : n_act + __popcll(~bal & below);
}

// Seeds for the active lanes of a wave from their NIN inputs f: active pairs
// are compacted to the first rows, so a wave with k active lanes runs
// ceil(k/16) pair blocks.  Called by all 64 lanes; u, v and the hit
// probability ph are written for active lanes only.
__device__ __forceinline__ void seed_wave(const bf16x8* __restrict__ tiles, const float* __restrict__ bias,
                                          bool active, const float f[NIN], float& u, float& v, float& ph,
                                          bf16x8 (*s_x)[4], float4* s_logit, int lane) {
  const unsigned long long bal = __ballot(active);
  const int n_act = __popcll(bal);
  const unsigned long long below = lane ? (~0ull >> (64 - lane)) : 0ull;
  const int row = active ? __popcll(bal & below) : n_act + __popcll(~bal & below);
  store_row(s_x, row, f, !active);
  wave_lds_sync();
  forward_wave<false>(tiles, bias, s_x, (n_act + 15) >> 4, lane, s_logit, nullptr);
  wave_lds_sync();
  if (active) {
    const float4 lg = s_logit[row];
    u = fast_sigmoid(lg.x);
    v = fast_sigmoid(lg.y);
    ph = fast_sigmoid(lg.z);
  }
  wave_lds_sync();  // the next call's writes come after these reads
}

}  // namespace mlp
}  // namespace rtn
