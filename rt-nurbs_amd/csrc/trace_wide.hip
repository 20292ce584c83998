// Fast trace forms over the 4-wide SoA BVH: the wide kernel (one ray per
// lane) and the persistent kernel (lanes refill from a global ray counter).
//
// Per lane: a short LDS stack of (node, entry t) entries ordered near-first;
// popped entries are re-tested against the current closest hit.  The wave runs
// in two phases (while-while): lanes without a pending leaf advance through
// the tree until each has one, then the leaf phase processes one patch per
// lane.  Newton is dispatched per degree class: the wave takes the class of
// its first pending lane (readfirstlane), runs a compile-time Newton<P,Q> for
// the lanes of that class, and repeats until no lane is pending, so no lane
// ever runs a runtime-degree evaluator.  Control nets stream row by row from
// L1/L2 (the rays of a wave mostly test the same patches).
//
// Arithmetic: explicit FMA and tensor-product Bernstein evaluation (rtn_fast.hpp);
// the seed is the reference's bilinear seed in its own operation order; the
// iteration keeps the plain method's convergence test, PCG perturbation on a
// singular Jacobian and clamping.  Options: exact rejection of primitives from
// their projected control points at leaf entry, the learned seed (network on
// the matrix cores for all pending pairs of the wave, plain seed as fallback),
// an exit for solves stuck against the patch boundary, and an FP64 correction
// step for hits whose Jacobian condition number exceeds a threshold (mixed
// precision).  Improving hits are written through to the output arrays, as the
// reference writes its record (intersect.cu:133-138).
#include <algorithm>

#include "mlp.hpp"
#include "rtn_fast.hpp"
#include "rtn_internal.hpp"
#include "trace_wide.hpp"

using namespace rtn;
using namespace rtn::fast;

namespace {

struct Cnt {  // per-ray work counters (counting builds only)
  uint32_t nc, ne, ncull, nbox, nleaf, npt, ovf, nfl, nfb, nskip;  // (FP64 steps are counted by k_fix64)
};

struct Lane {
  f3 o, d, rd;
  Plane2 rp;
  uint32_t r;
  float t_max;
  int best;
  bool fix;  // mixed precision: the closest hit so far is ill-conditioned
};

// Starts ray r on this lane; returns its root node (the ray's scene in a
// batch), or -1 for an empty secondary slot (recorded as a miss at once).
__device__ __forceinline__ int lane_start(Lane& L, const WArgs& A, uint32_t r) {
  if (A.ray_kind && A.ray_kind[r] < 0) {
    A.prim[r] = -1;
    return -1;
  }
  L.r = r;
  L.o = {A.ox[r], A.oy[r], A.oz[r]};
  L.d = {A.dx[r], A.dy[r], A.dz[r]};
  // finite reciprocals keep o*rd free of 0*inf (a zero component only widens
  // the slab interval, so the box test stays conservative)
  L.rd = {1.0f / (fabsf(L.d.x) > 1e-30f ? L.d.x : copysignf(1e-30f, L.d.x)),
          1.0f / (fabsf(L.d.y) > 1e-30f ? L.d.y : copysignf(1e-30f, L.d.y)),
          1.0f / (fabsf(L.d.z) > 1e-30f ? L.d.z : copysignf(1e-30f, L.d.z))};
  L.rp = planes(L.o, L.d);
  L.t_max = INFINITY;
  L.best = -1;
  L.fix = false;
  return A.scene_root ? A.scene_root[A.ray_scene[r]] : 0;
}

// The patch ray r does not test (the patch a secondary ray leaves), or -1;
// only called when A.skip_prim is set.  Kept in LDS, not in a register: the
// launches without it (primary rays) keep their register budget.
__device__ __forceinline__ int skip_patch_of(const WArgs& A, uint32_t r) {
  return A.skip_prim[A.skip_parent ? A.skip_parent[r] : r];
}

template <bool COUNT>
__device__ __forceinline__ void lane_finish(const Lane& L, const WArgs& A, const Cnt& c) {
  // a closest hit that wants the FP64 step is marked -2 - patch until the
  // correction pass (k_fix64) has run it
  A.prim[L.r] = L.fix ? -2 - L.best : L.best;
  if (COUNT && A.counters) {
    uint32_t* o = A.counters + 8 * (int64_t)L.r;
    o[0] = c.nc; o[1] = c.ne; o[2] = c.ncull; o[3] = c.nbox; o[4] = c.nleaf; o[5] = c.npt; o[6] = c.ovf; o[7] = c.nfl;
  }
}

// Expand one wide node: test the four child boxes (slab test with FMA,
// t = b*rd - o*rd) and push the hits far-to-near so the nearest is on top.
template <int WB, int WSTK, bool COUNT>
__device__ __forceinline__ void expand(Lane& L, Cnt& cn, const WideNode* __restrict__ nodes, int ni, int (*sc)[WB],
                                       float (*st)[WB], int tid, int& off) {
  const float4* base = reinterpret_cast<const float4*>(nodes + ni);
  const float4 lx = base[0], hx = base[1], ly = base[2], hy = base[3], lz = base[4], hz = base[5];
  const int4 ch = *reinterpret_cast<const int4*>(base + 6);
  const float lxa[4] = {lx.x, lx.y, lx.z, lx.w}, hxa[4] = {hx.x, hx.y, hx.z, hx.w};
  const float lya[4] = {ly.x, ly.y, ly.z, ly.w}, hya[4] = {hy.x, hy.y, hy.z, hy.w};
  const float lza[4] = {lz.x, lz.y, lz.z, lz.w}, hza[4] = {hz.x, hz.y, hz.z, hz.w};
  const int cha[4] = {ch.x, ch.y, ch.z, ch.w};
  const float ox = L.o.x * L.rd.x, oy = L.o.y * L.rd.y, oz = L.o.z * L.rd.z;
  float te[4];
  int code[4];
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const float tx0 = fmaf(lxa[k], L.rd.x, -ox), tx1 = fmaf(hxa[k], L.rd.x, -ox);
    const float ty0 = fmaf(lya[k], L.rd.y, -oy), ty1 = fmaf(hya[k], L.rd.y, -oy);
    const float tz0 = fmaf(lza[k], L.rd.z, -oz), tz1 = fmaf(hza[k], L.rd.z, -oz);
    const float tmin = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fmaxf(fminf(tz0, tz1), 0.0f));
    // a few ulps of slack on the exit distance keep the test conservative
    const float tmax = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fmaxf(tz0, tz1)) * 1.0000005f;
    // empty slots carry the sentinel child 0x7fffffff and are never entered
    const bool hit = tmin <= tmax && tmin < L.t_max && cha[k] != 0x7fffffff;
    te[k] = hit ? tmin : INFINITY;
    code[k] = cha[k];
  }
  if (COUNT) cn.nbox += 4;
#define RTN_CSWAP(a, b)                                \
  if (te[a] < te[b]) {                                 \
    float tt = te[a]; te[a] = te[b]; te[b] = tt;       \
    int cc = code[a]; code[a] = code[b]; code[b] = cc; \
  }
  RTN_CSWAP(0, 1) RTN_CSWAP(2, 3) RTN_CSWAP(0, 2) RTN_CSWAP(1, 3) RTN_CSWAP(1, 2)
#undef RTN_CSWAP
#pragma unroll
  for (int k = 0; k < 4; k++) {
    if (te[k] < INFINITY) {
      if (off < WSTK) {
        sc[off][tid] = code[k];
        st[off][tid] = te[k];
        off++;
      } else {
        cn.ovf = 1;
      }
    }
  }
}

// The fast forms keep each lane's closest-hit record (u, v, dpdu, dpdv; t is
// L.t_max) in LDS, one column per lane, and write it once when the ray is
// done, instead of writing it through on every improving hit (the packet
// form's 64 lanes finish together, so a chunk's stores are contiguous).  The
// one-ray and persistent forms with the learned seed write it through
// (record_global): their LDS already holds the stacks and the network's input
// rows, and 32 more bytes per lane would cost them resident blocks.
constexpr int kRecWords = 8;
__device__ __forceinline__ void record_global(Lane& L, const WArgs& A, int pi, float t, float u, float v, f3 Su,
                                              f3 Sv, bool fix) {
  L.t_max = t;
  L.best = pi;
  L.fix = fix;
  const int64_t r = L.r;
  if (A.t) A.t[r] = t;
  if (A.u) A.u[r] = u;
  if (A.v) A.v[r] = v;
  if (A.dpdu) { A.dpdu[3 * r] = Su.x; A.dpdu[3 * r + 1] = Su.y; A.dpdu[3 * r + 2] = Su.z; }
  if (A.dpdv) { A.dpdv[3 * r] = Sv.x; A.dpdv[3 * r + 1] = Sv.y; A.dpdv[3 * r + 2] = Sv.z; }
}

template <int W>
__device__ __forceinline__ void record_lds(Lane& L, float (*rec)[W], int lane, int pi, float t, float u, float v,
                                           f3 Su, f3 Sv, bool fix) {
  L.t_max = t;
  L.best = pi;
  L.fix = fix;
  rec[0][lane] = u;
  rec[1][lane] = v;
  rec[2][lane] = Su.x; rec[3][lane] = Su.y; rec[4][lane] = Su.z;
  rec[5][lane] = Sv.x; rec[6][lane] = Sv.y; rec[7][lane] = Sv.z;
}

template <int W>
__device__ __forceinline__ void write_record(const Lane& L, const WArgs& A, const float (*rec)[W], int lane) {
  if (L.best < 0) return;
  const int64_t r = L.r;
  if (A.t) A.t[r] = L.t_max;
  if (A.u) A.u[r] = rec[0][lane];
  if (A.v) A.v[r] = rec[1][lane];
  if (A.dpdu) { A.dpdu[3 * r] = rec[2][lane]; A.dpdu[3 * r + 1] = rec[3][lane]; A.dpdu[3 * r + 2] = rec[4][lane]; }
  if (A.dpdv) { A.dpdv[3 * r] = rec[5][lane]; A.dpdv[3 * r + 1] = rec[6][lane]; A.dpdv[3 * r + 2] = rec[7][lane]; }
}

// Result of one patch test.
struct LeafOut {
  float t, u, v;
  f3 S, Su, Sv;
  int evals;     // surface evaluations (both attempts when a learned seed falls back)
  int flags;     // 1 hit, 8 learned fallback, 16 FP64 step wanted (ill-conditioned hit)
};

struct LeafIn {
  Plane2 rp;
  f3 o, d;
  float gu, gv, pu, pv;
  uint32_t seed;
  int max_iter, learned_iters, fallback_iters, mixed, learned, stall;
  float thr_det, thr_res, cond_thr, t_tol;
};

// Leaf work for one patch with compile-time row length Q+1 and runtime row
// count P+1: Newton from the seed (plain-seed fallback for the learned form),
// flagging an FP64 correction.
template <int Q, class NP = const float4*>
__device__ __forceinline__ LeafOut leaf_q(NP net, int P, const LeafIn& in) {
  LeafOut o;
  o.flags = 0;
  o.evals = 0;
  // attempt 0 from the given seed; for the learned form a failed attempt is
  // repeated from the plain seed
  Hit h;
  float su = in.gu, sv = in.gv;
#pragma unroll 1
  for (int attempt = 0;; attempt++) {
    h = newton_q<Q>(net, P, in.rp, su, sv, in.thr_det, in.thr_res, in.seed,
                    !in.learned ? in.max_iter : attempt == 0 ? in.learned_iters : in.fallback_iters, in.stall != 0);
    o.evals += h.evals;
    if (h.hit || !in.learned || attempt == 1) break;
    o.flags |= 8;
    su = in.pu;
    sv = in.pv;
  }
  if (!h.hit) return o;
  const float t = fmaf(h.S.x - in.o.x, in.d.x, fmaf(h.S.y - in.o.y, in.d.y, (h.S.z - in.o.z) * in.d.z));
  // FP64 step wanted (flag 16): an ill-conditioned hit (cond(J) above the
  // threshold), or one whose predicted t error (the change one more Newton
  // step would make) exceeds t_tol relative.  If it ends as the ray's closest
  // hit, the correction pass k_fix64 runs the FP64 step on it after the trace,
  // so the FP64 code and its registers stay out of the trace kernels.
  if (in.mixed && (cond2(h.m0, h.m1, h.m2, h.m3) > in.cond_thr || predicted_dt(h, in.d) > in.t_tol * t))
    o.flags |= 16;
  o.flags |= 1;
  o.t = t;
  o.u = h.u;
  o.v = h.v;
  o.S = h.S;
  o.Su = h.Su;
  o.Sv = h.Sv;
  return o;
}

// Degree classes: the row length Q+1 is compiled (q = 1..5 -> class 0..4);
// the row count P+1 is a runtime loop.
__device__ __forceinline__ int degree_class(int p, int q) { return q - 1; }

template <class NP = const float4*>
__device__ __forceinline__ LeafOut leaf_dispatch(int cls, NP net, int p, const LeafIn& in) {
  switch (cls) {  // wave-uniform: a scalar branch
    case 0: return leaf_q<1>(net, p, in);
    case 1: return leaf_q<2>(net, p, in);
    case 2: return leaf_q<3>(net, p, in);
    case 3: return leaf_q<4>(net, p, in);
    default: return leaf_q<5>(net, p, in);
  }
}

// The learned form's use of the network's hit probability ph on a lane's
// pair (test: the lane solves the pair): below hit_skip the pair is not
// solved (a skip); below hit_gate it starts from the plain seed with the full
// iteration count, as without the network.
__device__ __forceinline__ void gate_pair(const WArgs& A, float ph, bool& test, bool& lrn, LeafIn& in,
                                          uint32_t& nskip) {
  if (!test) return;
  if (ph < A.hit_skip) {
    test = false;
    nskip++;
  } else if (ph < A.hit_gate) {
    lrn = false;
    in.gu = in.pu;
    in.gv = in.pv;
  }
}


// The kernel's argument block read again from the kernel-argument segment at
// the point of use: a scalar load through an opaque pointer, which keeps the
// load there.  The trace kernels hold more values in scalar registers than
// the 106 a wave has (control-net rows, node boxes, branch masks); arguments
// kept live across their loops were spilled to vector-register lanes and read
// back with one VALU instruction per register at every use.  Valid in the
// kernels whose only argument is their WArgs, passed by value (it starts the
// segment).
__device__ __forceinline__ const WArgs& args_at_use() {
  auto p = (const __attribute__((address_space(4))) WArgs*)__builtin_amdgcn_kernarg_segment_ptr();
  asm volatile("" : "+s"(p));
  return *(const WArgs*)p;
}

template <int WB, int WSTK, bool PERSISTENT, bool COUNT, bool LEARNED>
__device__ __forceinline__ void trace_wide_body(const WArgs& A) {
  __shared__ int s_code[WSTK][WB];
  __shared__ float s_te[WSTK][WB];
  __shared__ int s_skip[WB];  // per lane, skip_patch_of its ray
  __shared__ float s_rec[LEARNED ? 1 : kRecWords][LEARNED ? 1 : WB];  // per lane, its closest hit's record
  // learned form: per-wave feature rows of the seed network (24 BF16 per row)
  __shared__ mlp::bf16x8 s_x[LEARNED ? WB : 1][3];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  Lane L;
  Cnt cn{};
  bool has_ray = false;   // a ray is in flight on this lane
  int off = 0;            // stack depth
  int cur = -1;           // wide node to expand, or -1
  int cand_first = 0, cand_left = 0;  // primitives of the current leaf still to test
  bool exhausted;         // no more rays to hand out
  int queue = blockIdx.x % kQueues, queues_left = kQueues;  // persistent: the wave's work queue
  // work totals of the lane: all twelve when counting, otherwise only the
  // stack-overflow count (total 6), so no idle accumulators occupy registers
  constexpr int NACC = COUNT ? 12 : 1;
  unsigned long long acc[NACC] = {};
  exhausted = false;
  // One loop shape for both forms (it also keeps the one-ray form's register
  // allocation as small as the persistent one's).
  while (true) {
    if (!PERSISTENT && !exhausted) {  // the lane's one ray, taken on the first pass
      const WArgs& A = args_at_use();
      exhausted = true;
      const int64_t r = (int64_t)blockIdx.x * WB + tid;
      if (r < A.n_rays) {
        cur = lane_start(L, A, (uint32_t)r);
        if (A.skip_prim) s_skip[tid] = skip_patch_of(A, (uint32_t)r);
        has_ray = cur >= 0;
      }
    }
    if (PERSISTENT && !exhausted) {
      const WArgs& A = args_at_use();
      // lanes without a ray take consecutive positions of the wave's work
      // queue, one atomic per wave; a wave takes new rays once `refill_min` of
      // its lanes are idle (or all of them)
      const unsigned long long need = __ballot(!has_ray);
      if (need && (__popcll(need) >= A.refill_min || need == ~0ull)) {
        const int cnt = __popcll(need);
        const int leader = __ffsll((long long)need) - 1;
        unsigned int base = 0;
        if (lane == leader) base = atomicAdd(A.work + kQueueStride * queue, (unsigned)cnt);
        base = __shfl(base, leader);
        const int q = queue;
        if (queue_ray(q, base + (unsigned)cnt - 1u) >= A.n_rays) {  // this queue has run past the last ray
          queue = (queue + 1) % kQueues;
          if (--queues_left == 0) exhausted = true;
        }
        if (!has_ray) {
          const unsigned rank = __popcll(need & ((1ull << lane) - 1));
          const int64_t r = queue_ray(q, base + rank);
          if (r < A.n_rays) {
            cur = lane_start(L, A, (uint32_t)r);
            if (A.skip_prim) s_skip[tid] = skip_patch_of(A, (uint32_t)r);
            if (COUNT) cn = Cnt{};
            has_ray = cur >= 0;
            off = 0;
            cand_left = 0;
          }
        }
      }
    }
    // ---- traversal: advance every lane that has no pending leaf.  The
    // leaf's candidates are checked one at a time before Newton: the patch a
    // secondary ray leaves is skipped, and with hull_cull each candidate is
    // first tested against the oriented box, then the projected control points.
    while (has_ray) {
      if (cand_left > 0) {
        const WArgs& A = args_at_use();
        bool drop = A.skip_prim && cand_first == s_skip[tid];
        if (!drop && A.hull_cull) {
          const int pqv = A.pq[cand_first];
          if (pqv & 0x10000) {
            drop = obb_culled(A.obb + 4 * (int64_t)cand_first, L.o, L.d, A.cull_margin);
            if (!drop) {
              const int np = ((pqv & 0xff) + 1) * (((pqv >> 8) & 0xff) + 1);
              drop = hull2d_culled(A.cp + A.cp_off[cand_first], np, L.rp, A.cull_margin, A.hull_cull == 2);
            }
            if (COUNT && drop) {
              cn.ncull++;
              cn.npt++;
            }
          }
        }
        if (!drop) break;
        cand_first++;
        cand_left--;
        continue;
      }
      if (cur >= 0) {
        expand<WB, WSTK, COUNT>(L, cn, args_at_use().nodes, cur, s_code, s_te, tid, off);
        cur = -1;
        continue;
      }
      if (off == 0) {  // traversal finished: the ray is done
        const WArgs& A = args_at_use();
        if constexpr (!LEARNED) write_record<WB>(L, A, reinterpret_cast<const float(*)[WB]>(&s_rec[0][0]), tid);
        lane_finish<COUNT>(L, A, cn);
        if constexpr (COUNT) {
          acc[0] += cn.nc; acc[1] += cn.ne; acc[3] += cn.nbox; acc[4] += cn.nleaf; acc[5] += cn.npt;
          acc[6] += cn.ovf; acc[7] += cn.nfl; acc[9] += cn.ncull; acc[10] += cn.nfb; acc[11] += cn.nskip;
        } else if (cn.ovf) {
          acc[0] += 1;
          cn.ovf = 0;
        }
        has_ray = false;
        break;
      }
      off--;
      const int code = s_code[off][tid];
      const float te = s_te[off][tid];
      if (te >= L.t_max) continue;
      if (code < 0) {
        cand_first = code & 0xffffff;
        cand_left = ((code >> 24) & 0x7f) + 1;
        if (COUNT) cn.nleaf++;
      } else {
        cur = code;
      }
    }
    const bool pending = cand_left > 0;
    if (!__ballot(pending)) {
      if (exhausted && !__ballot(has_ray)) break;
      continue;
    }
    // ---- leaf phase: one primitive per pending lane, Newton by degree class
    const WArgs& A = args_at_use();
    int pi = 0, p = 0, q = 0, cls = -1;
    float gu = 0.5f, gv = 0.5f;
    float4 c0 = make_float4(0.f, 0.f, 0.f, 0.f), c1 = c0, c2 = c0;
    if (pending) {
      pi = cand_first;
      cand_first++;
      cand_left--;
      if (COUNT) cn.npt++;
      const int pqv = A.pq[pi];
      p = pqv & 0xff;
      q = (pqv >> 8) & 0xff;
      cls = degree_class(p, q);
      if (A.mixed) {  // the FP64 reference's seed
        seed64(A.seed64 + (int64_t)kSeedTermWords * pi, L.o, L.d, gu, gv);
      } else {  // the reference's bilinear seed on the dehomogenised corners (bezier_patch.h:38-44, :273-327)
        c0 = A.corners[3 * pi];
        c1 = A.corners[3 * pi + 1];
        c2 = A.corners[3 * pi + 2];
        ex::bilinear_seed<float>({c0.x, c0.y, c0.z}, {c0.w, c1.x, c1.y}, {c1.z, c1.w, c2.x}, {c2.y, c2.z, c2.w},
                                 {L.o.x, L.o.y, L.o.z}, {L.d.x, L.d.y, L.d.z}, gu, gv);
      }
    }
    LeafIn in;
    in.rp = L.rp;
    in.o = L.o;
    in.d = L.d;
    in.gu = gu;
    in.gv = gv;
    in.pu = gu;
    in.pv = gv;
    bool test = pending;  // the pair is solved (the network's hit probability may skip it)
    // the pair starts from the network's seed; with fewer than learned_min_lanes
    // pending lanes the wave keeps the plain seeds
    const bool net_pass = LEARNED && __popcll(__ballot(pending)) >= A.learned_min_lanes;
    bool lrn = net_pass;
    if (net_pass) {  // network seed for every pending pair of the wave (matrix cores), plain seed kept as fallback
      // with seed sharing, a lane whose quad leader (lane & ~3) tests the same
      // patch takes the leader's seed instead of an evaluation of its own
      const int lead = lane & ~3;
      const int lead_pi = __shfl(pending ? pi : -1, lead);
      const bool own = pending && (!A.seed_share || lane == lead || lead_pi != pi);
      auto rows = reinterpret_cast<mlp::bf16x8(*)[3]>(&s_x[tid & ~63][0]);
      int n_act;
      const int row = mlp::seed_row(own, lane, n_act);
      if (own) {
        mlp::store_features(rows, row, A.mlp_kind, L.o.x, L.o.y, L.o.z, L.d.x, L.d.y, L.d.z, L.rp.n1.x, L.rp.n1.y,
                            L.rp.n1.z, L.rp.n2.x, L.rp.n2.y, L.rp.n2.z, A.cp + A.cp_off[pi], p, q,
                            A.bounds + 6 * (int64_t)pi);
      } else {
        const mlp::bf16x8 z = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
        for (int k = 0; k < 3; k++) rows[row][k] = z;
      }
      float ph = 1.0f;
      mlp::seed_wave_rows<false>(A.mlp_tiles, A.mlp_bias, own, row, n_act, in.gu, in.gv, ph, rows, lane);
      if (A.seed_share) {
        const float lu = __shfl(in.gu, lead), lv = __shfl(in.gv, lead), lp = __shfl(ph, lead);
        if (pending && !own) {
          in.gu = lu;
          in.gv = lv;
          ph = lp;
        }
      }
      gate_pair(A, ph, test, lrn, in, cn.nskip);
    }
    in.seed = L.r + A.seed_base;
    in.max_iter = A.max_iter;
    in.learned_iters = A.learned_iters;
    in.fallback_iters = A.fallback_iters;
    in.mixed = A.mixed;
    in.stall = A.stall_exit;
    in.learned = lrn;
    in.thr_det = A.thr_det;
    in.thr_res = A.thr_res;
    in.cond_thr = A.cond_thr;
    in.t_tol = A.t_tol;
    const float4* net = A.cp + A.cp_off[pi];
    LeafOut res;
    res.flags = 0;
    unsigned long long todo = __ballot(test);
    while (todo) {
      const int leader = __ffsll((long long)todo) - 1;
      const int c = __shfl(cls, leader);
      const bool mine = test && cls == c;
      if (mine) res = leaf_dispatch(c, net, p, in);
      todo &= ~__ballot(mine);
    }
    if (test) {
      if (COUNT) {
        cn.nc++;
        cn.ne += res.evals;
        cn.nfl += res.evals * ex::eval_flops(p, q);
        if (res.flags & 8) cn.nfb++;
      }
      const bool conv = (res.flags & 1) && 0.0f < res.t;
      if (A.pair_count && (conv || A.pair_all)) {
        const unsigned long long k = atomicAdd(A.pair_count, 1ull);
        if ((int64_t)k < A.pair_cap) {
          A.pair_ray[k] = (int)L.r;
          A.pair_patch[k] = pi;
          A.pair_uv[k] = conv ? make_float2(res.u, res.v) : make_float2(-1.0f, -1.0f);
        }
      }
      if ((res.flags & 1) && 0.0f < res.t && res.t < L.t_max) {
        // an occlusion query's answer is hit or miss, which the FP64 step
        // does not change: its hit is not flagged for the correction pass
        const bool occl = A.any_hit || (A.ray_kind && A.ray_kind[L.r] == 0);
        const bool fix = (res.flags & 16) != 0 && !occl;
        if constexpr (LEARNED) record_global(L, A, pi, res.t, res.u, res.v, res.Su, res.Sv, fix);
        else record_lds<WB>(L, reinterpret_cast<float(*)[WB]>(&s_rec[0][0]), tid, pi, res.t, res.u, res.v, res.Su,
                            res.Sv, fix);
        if (occl) {  // occlusion query: the first hit ends the ray
          off = 0;
          cand_left = 0;
          cur = -1;
        }
      }
    }
  }
  if (A.totals) {  // every lane of the wave arrives here together
#pragma unroll
    for (int k = 0; k < NACC; k++) {
      unsigned long long x = acc[k];
      for (int sft = 32; sft > 0; sft >>= 1) x += __shfl_xor(x, sft);
      if (lane == 0 && x) atomicAdd(A.totals + (COUNT ? k : 6), x);
    }
  }
}

// The arguments travel by value in the kernel-argument segment: pointers read
// from there are known to address global memory, so every load and store of
// the kernel is a global (not flat) access and LDS waits never wait on them.
template <int WB, int WSTK, bool PERSISTENT, bool COUNT, bool LEARNED>
__global__ void __launch_bounds__(WB) k_trace_wide(const WArgs A) {
  trace_wide_body<WB, WSTK, PERSISTENT, COUNT, LEARNED>(A);
}

// The learned forms held to three waves per SIMD (168 registers per lane
// with the matrix accumulators, a few bytes of spill; left to itself the
// compiler takes about 200 and runs two).
template <int WB, int WSTK, bool PERSISTENT, bool COUNT>
__global__ void __launch_bounds__(WB) __attribute__((amdgpu_waves_per_eu(3, 3)))
k_trace_wide_learned(const WArgs A) {
  trace_wide_body<WB, WSTK, PERSISTENT, COUNT, true>(A);
}

// ------------------------------------------------------------ packet form
// One wave traverses the 4-wide BVH for its 64 rays together.  The wave
// visits a node when any of its lanes enters one of the node's child boxes;
// each lane tests the child boxes against its own ray and closest hit, so a
// lane enters exactly the boxes it would enter alone.  The stack is the
// wave's: entry i lives in lane i of one register (a select and readlane), and
// holds wide-node indices or, for a leaf, its parent node and slot, so that on
// the pop every lane re-tests the leaf box against its closest hit exactly as
// the one-ray form tests the entry distance it pushed.  A leaf's patches are
// then tested one at a time by the lanes that entered it: node, box and
// control-point addresses are the same on every lane (one cache line per
// wave), and every Newton solve of a pass has the patch's degree, so there is
// no degree-class divergence.  Children are ordered near-first by the entry
// distance of the first lane that enters each.
//
// Per (ray, patch) the arithmetic is the one-ray form's; a ray tests every
// patch whose box it enters before its closest hit, so the closest hit is the
// one-ray form's except where two patches give the same t (the first tested
// is kept).  Plain or learned seed; single scenes (no multi-scene batches).
__device__ __forceinline__ int rl_int(int x, int l) { return __builtin_amdgcn_readlane(x, l); }
__device__ __forceinline__ float rl_float(float x, int l) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), l)); }

// Component k (a compile-time index after unrolling) of a float4.
__device__ __forceinline__ float f4c(const float4& v, int k) { return k == 0 ? v.x : k == 1 ? v.y : k == 2 ? v.z : v.w; }

// The slab interval of the lane's ray through the box [lx,hx] x [ly,hy] x
// [lz,hz] (child_box's arithmetic, on values already in registers); the ray
// enters the box when tmin <= tmax and tmin < L.t_max.
__device__ __forceinline__ void slab_box(float lx, float hx, float ly, float hy, float lz, float hz, const Lane& L,
                                         float& tmin, float& tmax) {
  const float ox = L.o.x * L.rd.x, oy = L.o.y * L.rd.y, oz = L.o.z * L.rd.z;
  const float tx0 = fmaf(lx, L.rd.x, -ox), tx1 = fmaf(hx, L.rd.x, -ox);
  const float ty0 = fmaf(ly, L.rd.y, -oy), ty1 = fmaf(hy, L.rd.y, -oy);
  const float tz0 = fmaf(lz, L.rd.z, -oz), tz1 = fmaf(hz, L.rd.z, -oz);
  tmin = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fmaxf(fminf(tz0, tz1), 0.0f));
  tmax = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fmaxf(tz0, tz1)) * 1.0000005f;
}

// Entry and exit distances of the lane's ray through child slot k of a node
// (the one-ray form's slab test, expand()).
__device__ __forceinline__ bool child_box(cptr<float> f, int k, const Lane& L, float& tmin) {
  const float ox = L.o.x * L.rd.x, oy = L.o.y * L.rd.y, oz = L.o.z * L.rd.z;
  const float tx0 = fmaf(f[k], L.rd.x, -ox), tx1 = fmaf(f[4 + k], L.rd.x, -ox);
  const float ty0 = fmaf(f[8 + k], L.rd.y, -oy), ty1 = fmaf(f[12 + k], L.rd.y, -oy);
  const float tz0 = fmaf(f[16 + k], L.rd.z, -oz), tz1 = fmaf(f[20 + k], L.rd.z, -oz);
  tmin = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fmaxf(fminf(tz0, tz1), 0.0f));
  const float tmax = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fmaxf(tz0, tz1)) * 1.0000005f;
  return tmin <= tmax && tmin < L.t_max;
}

constexpr int PKT_STACK = 64;  // one entry per lane of the stack register

// The 64 rays r0 .. r0+63 (those below n_rays) traced as one packet by the
// calling wave; work totals accumulate into acc.
// The inputs of a patch test of the packet form: the ray, the plain seed (the
// FP64 reference's seed with mixed precision) and the method's parameters.
// The plain seed of the lane's ray on patch pi (the FP64 reference's seed with
// mixed precision); the learned kernel computes it again for a fallback
// attempt instead of keeping it live across the network pass.
__device__ __forceinline__ void packet_plain_seed(const Lane& L, const WArgs& A, int pi, float& su, float& sv) {
  if (A.mixed) {
    seed64(as_const(A.seed64 + (int64_t)kSeedTermWords * pi), L.o, L.d, su, sv);
  } else {
    const cptr<float4> cn4 = as_const(A.corners + 3 * (int64_t)pi);
    const float4 c0 = ld4(cn4, 0), c1 = ld4(cn4, 1), c2 = ld4(cn4, 2);
    ex::bilinear_seed<float>({c0.x, c0.y, c0.z}, {c0.w, c1.x, c1.y}, {c1.z, c1.w, c2.x}, {c2.y, c2.z, c2.w},
                             {L.o.x, L.o.y, L.o.z}, {L.d.x, L.d.y, L.d.z}, su, sv);
  }
}

__device__ __forceinline__ void packet_leaf_in(LeafIn& in, const Lane& L, const WArgs& A, int pi) {
  in.rp = L.rp;
  in.o = L.o;
  in.d = L.d;
  packet_plain_seed(L, A, pi, in.gu, in.gv);
  in.seed = L.r + A.seed_base;
  in.max_iter = A.max_iter;
  in.learned_iters = A.learned_iters;
  in.fallback_iters = A.fallback_iters;
  in.mixed = A.mixed;
  in.stall = A.stall_exit;
  in.thr_det = A.thr_det;
  in.thr_res = A.thr_res;
  in.cond_thr = A.cond_thr;
  in.t_tol = A.t_tol;
}

template <bool COUNT, bool LEARNED, int NACC, class Acc>
__device__ __forceinline__ void packet_chunk(const WArgs& A, int64_t r0, Acc (&acc)[NACC],
                                             mlp::bf16x8 (*s_x)[3], int* s_skip, float (*s_rec)[64]) {
  const int lane = threadIdx.x & 63;
  Lane L;
  Cnt cn{};
  int root = -1;
  {
    const WArgs& A = args_at_use();
    const int64_t r = r0 + lane;
    if (r < A.n_rays) root = lane_start(L, A, (uint32_t)r);
    if (A.skip_prim) s_skip[lane] = root >= 0 ? skip_patch_of(A, (uint32_t)r) : -1;
  }
  bool pending = root >= 0;  // rays whose traversal has not run yet
  const bool had_ray = pending;
  // one packet per distinct scene root among the wave's rays (in a batch the
  // rays of one scene are contiguous, so a wave almost always has one)
  while (true) {
  const unsigned long long todo = __ballot(pending);
  if (!todo) break;
  int e = rl_int(root, __ffsll((long long)todo) - 1);  // entry being processed: wide node (>= 0) or
                                                       // (parent << 2 | slot) | sign bit for a leaf
  bool active = pending && root == e;
  pending = pending && !active;
  int stk = 0;  // the wave's stack: entry i in lane i
  int sp = 0;   // wave-uniform depth
  while (__ballot(active)) {
    if (e >= 0) {
      const WArgs& A = args_at_use();
      // the whole node in one batch of scalar loads (no load behind a per-child
      // branch), then the four slab tests on every lane, masked afterwards
      const cptr<float4> nd = as_const(reinterpret_cast<const float4*>(A.nodes + e));
      const float4 bx[6] = {ld4(nd, 0), ld4(nd, 1), ld4(nd, 2), ld4(nd, 3), ld4(nd, 4), ld4(nd, 5)};
      const cptr<int> chp = as_const(reinterpret_cast<const int*>(A.nodes + e) + 24);
      const int cha[4] = {chp[0], chp[1], chp[2], chp[3]};
      // entry distances as their bit patterns: every key is +0 .. +inf, where the
      // unsigned order of the bits is the order of the values, so the sorting
      // network runs on the scalar unit
      uint32_t key[4];
      int code[4];
      const unsigned long long act_mask = __ballot(active);
#pragma unroll
      for (int k = 0; k < 4; k++) {
        float tmin, tmax;
        slab_box(f4c(bx[0], k), f4c(bx[1], k), f4c(bx[2], k), f4c(bx[3], k), f4c(bx[4], k), f4c(bx[5], k), L, tmin,
                 tmax);
        // the lanes that enter the child: a ballot per compare (each a compare's own
        // lane mask) and scalar ANDs, no per-lane boolean to materialise
        const unsigned long long slot_mask = cha[k] != 0x7fffffff ? ~0ull : 0ull;  // no branch: the compiler
        // would split the slab arithmetic around one
        const unsigned long long m = __ballot(tmin <= tmax) & __ballot(tmin < L.t_max) & act_mask & slot_mask;
        key[k] = 0x7f800000u;
        if (m) key[k] = (uint32_t)rl_int(__float_as_int(tmin), __ffsll((long long)m) - 1) & 0x7fffffffu;
        // leaves are pushed as their parent and slot, re-tested on the pop
        code[k] = cha[k] >= 0 ? cha[k] : (int)(0x80000000u | ((uint32_t)e << 2) | (uint32_t)k);
      }
      if (COUNT && active) cn.nbox += 4;
#define RTN_CSWAP(a, b)                                      \
  if (key[a] < key[b]) {                                     \
    uint32_t tt = key[a]; key[a] = key[b]; key[b] = tt;      \
    int cc = code[a]; code[a] = code[b]; code[b] = cc;       \
  }
      RTN_CSWAP(0, 1) RTN_CSWAP(2, 3) RTN_CSWAP(0, 2) RTN_CSWAP(1, 3) RTN_CSWAP(1, 2)
#undef RTN_CSWAP
      // descending keys: push far to near so the nearest is on top
#pragma unroll
      for (int k = 0; k < 4; k++) {
        if (key[k] < 0x7f800000u) {
          if (sp < PKT_STACK) {
            stk = lane == sp ? code[k] : stk;
            sp++;
          } else if (active) {
            cn.ovf = 1;
          }
        }
      }
    } else {
      const int parent = (e & 0x7fffffff) >> 2, slot = e & 3;
      const cptr<float> base = as_const(reinterpret_cast<const float*>(A.nodes + parent));
      const int leaf = as_const(reinterpret_cast<const int*>(A.nodes + parent) + 24)[slot];
      float tmin;
      bool in_leaf = active & child_box(base, slot, L, tmin);  // no branch around the box's loads
      if (COUNT && in_leaf) cn.nleaf++;
      const int first = leaf & 0xffffff, count = ((leaf >> 24) & 0x7f) + 1;
#pragma unroll 1
      for (int pi = first; pi < first + count && __ballot(in_leaf); pi++) {
        const WArgs& A = args_at_use();
        bool act = in_leaf && (!A.skip_prim || pi != s_skip[lane]);  // the patch the ray leaves is not tested
        const int pqv = as_const(A.pq)[pi];
        const cptr<float4> net = as_const(A.cp + as_const(A.cp_off)[pi]);
        if (act && A.hull_cull && (pqv & 0x10000)) {
          bool culled = obb_culled(as_const(A.obb + 4 * (int64_t)pi), L.o, L.d, A.cull_margin);
          if (!culled) {
            const int np = ((pqv & 0xff) + 1) * (((pqv >> 8) & 0xff) + 1);
            culled = hull2d_culled(net, np, L.rp, A.cull_margin, A.hull_cull == 2);
          }
          if (culled) {
            act = false;
            if (COUNT) {
              cn.ncull++;
              cn.npt++;
            }
          }
        }
        if (!__ballot(act)) continue;
        const int p = pqv & 0xff, q = (pqv >> 8) & 0xff;
        LeafOut res;
        res.flags = 0;
        // leaf inputs: the learned kernel sets them on every lane (no per-lane merges
        // of copies that live across the network pass: no spills at four waves),
        // the plain kernel on the lanes that test the patch (fewer spills at five)
        LeafIn in;
        if (LEARNED || act) packet_leaf_in(in, L, A, pi);        // the pair starts from the network's seed; with fewer than learned_min_lanes
        // pairs on the patch the wave keeps the plain seeds
        const bool net_pass = LEARNED && __popcll(__ballot(act)) >= A.learned_min_lanes;
        bool lrn = net_pass;
        if (net_pass) {  // network seeds of the wave's pairs on this patch (matrix cores), plain seed as fallback
          const WArgs& A = args_at_use();
          // with seed sharing every lane of an aligned quad whose first lane tests the patch takes that lane's seed
          const int lead = lane & ~3;
          const bool lead_act = __shfl((int)act, lead) != 0;
          const bool own = act && (!A.seed_share || lane == lead || !lead_act);
          int n_act;
          const int row = mlp::seed_row(own, lane, n_act);
          if (own) {
            mlp::store_features(s_x, row, A.mlp_kind, L.o.x, L.o.y, L.o.z, L.d.x, L.d.y, L.d.z, L.rp.n1.x,
                                L.rp.n1.y, L.rp.n1.z, L.rp.n2.x, L.rp.n2.y, L.rp.n2.z, net, p, q,
                                as_const(A.bounds + 6 * (int64_t)pi));
          } else {
            const mlp::bf16x8 z = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
            for (int k = 0; k < 3; k++) s_x[row][k] = z;
          }
          float nu = 0.0f, nv = 0.0f, ph = 1.0f;
          mlp::seed_wave_rows<true>(A.mlp_tiles, A.mlp_bias, own, row, n_act, nu, nv, ph, s_x, lane);
          if (A.seed_share) {
            const float lu = __shfl(nu, lead), lv = __shfl(nv, lead), lp = __shfl(ph, lead);
            if (act && !own) {
              nu = lu;
              nv = lv;
              ph = lp;
            }
          }
          // the hit gate (gate_pair's rule): the plain seed stays below hit_gate
          if (act) {
            if (ph < A.hit_skip) {
              act = false;
              cn.nskip++;
            } else if (ph < A.hit_gate) {
              lrn = false;
            } else {
              in.gu = nu;
              in.gv = nv;
            }
          }
        }
        if constexpr (!LEARNED) {
          if (act) {
            if (COUNT) cn.npt++;
            in.learned = 0;
            res = leaf_dispatch(degree_class(p, q), net, p, in);  // the patch's degree: a wave-uniform branch
            if (COUNT) {
              cn.nc++;
              cn.ne += res.evals;
              cn.nfl += res.evals * ex::eval_flops(p, q);
            }
            if ((res.flags & 1) && 0.0f < res.t && res.t < L.t_max) {
              // an occlusion query takes no FP64 step (its answer is hit or miss)
              const bool occl = A.any_hit || (A.ray_kind && A.ray_kind[L.r] == 0);
              record_lds<64>(L, s_rec, lane, pi, res.t, res.u, res.v, res.Su, res.Sv, (res.flags & 16) != 0 && !occl);
              if (occl) {  // occlusion query: the first hit ends the ray
                active = false;
                in_leaf = false;
              }
            }
          }
        } else {
          if (COUNT && act) {
            cn.npt++;
            cn.nc++;
          }
          in.learned = 0;
          // attempt 0 from the lane's seed; a learned seed whose Newton does not
          // converge is solved again from the plain seed (attempt 1).  Each
          // attempt records its own hit, so no result of attempt 0 stays live
          // (per wave) while other lanes run attempt 1: the kernel keeps four
          // waves per SIMD
          bool solve = act;
          if (lrn) in.max_iter = A.learned_iters;
#pragma unroll 1
          for (int attempt = 0; __ballot(solve); attempt++) {
            if (solve) {
              res = leaf_dispatch(degree_class(p, q), net, p, in);  // the patch's degree: a wave-uniform branch
              if (COUNT) {
                cn.ne += res.evals;
                cn.nfl += res.evals * ex::eval_flops(p, q);
              }
              if ((res.flags & 1) && 0.0f < res.t && res.t < L.t_max) {
                const bool occl = A.any_hit || (A.ray_kind && A.ray_kind[L.r] == 0);
                record_lds<64>(L, s_rec, lane, pi, res.t, res.u, res.v, res.Su, res.Sv,
                               (res.flags & 16) != 0 && !occl);
                if (occl) {
                  active = false;
                  in_leaf = false;
                }
              }
              solve = lrn && attempt == 0 && !(res.flags & 1);
              if (solve) {
                if (COUNT) cn.nfb++;
                packet_plain_seed(L, A, pi, in.gu, in.gv);
                in.max_iter = A.fallback_iters;
              }
            }
          }
        }
      }
    }
    if (sp == 0) break;
    sp--;
    e = rl_int(stk, sp);
  }
  }
  if (had_ray) {
    const WArgs& A = args_at_use();
    write_record<64>(L, A, s_rec, lane);
    lane_finish<COUNT>(L, A, cn);
    if constexpr (COUNT) {
      acc[0] += cn.nc; acc[1] += cn.ne; acc[3] += cn.nbox; acc[4] += cn.nleaf; acc[5] += cn.npt;
      acc[6] += cn.ovf; acc[7] += cn.nfl; acc[9] += cn.ncull; acc[10] += cn.nfb; acc[11] += cn.nskip;
    } else if (cn.ovf) {
      acc[0] += 1;
    }
  }
}

// One packet per single-wave block, or (persistent) waves that take 64-ray
// chunks from a global counter until the rays run out: a wave finishes its
// whole chunk before it takes the next, so chunks keep the ray layout's
// coherence.
template <bool PERSISTENT, bool COUNT, bool LEARNED>
__device__ __forceinline__ void packet_body(const WArgs& A) {
  __shared__ mlp::bf16x8 s_x[LEARNED ? 64 : 1][3];
  __shared__ int s_skip[64];  // per lane, skip_patch_of its ray
  __shared__ float s_rec[kRecWords][64];  // per lane, its closest hit's record
  const int lane = threadIdx.x & 63;
  constexpr int NACC = COUNT ? 12 : 1;
  // without counting only the overflowed rays are summed: at most one per lane
  // and chunk, so 32 bits per lane hold it
  using Acc = std::conditional_t<COUNT, unsigned long long, uint32_t>;
  Acc acc[NACC] = {};
  if (!PERSISTENT) {
    packet_chunk<COUNT, LEARNED>(A, (int64_t)blockIdx.x * 64, acc, s_x, s_skip, s_rec);
  } else {
    // the work queues hand out 64-ray chunks (one position per chunk here)
    int queue = blockIdx.x % kQueues;
    for (int queues_left = kQueues; queues_left > 0;) {
      const WArgs& K = args_at_use();
      unsigned int idx = 0;
      if (lane == 0) idx = atomicAdd(K.work + kQueueStride * queue, 1u);
      idx = __builtin_amdgcn_readfirstlane(idx);  // wave-uniform: a scalar register
      const int64_t c = queue + (int64_t)kQueues * idx;
      if (c >= (K.n_rays + 63) / 64) {
        queue = (queue + 1) % kQueues;
        queues_left--;
        continue;
      }
      packet_chunk<COUNT, LEARNED>(A, c * 64, acc, s_x, s_skip, s_rec);
    }
  }
  if (A.totals) {  // every lane of the wave arrives here together
#pragma unroll
    for (int k = 0; k < NACC; k++) {
      unsigned long long x = acc[k];
      for (int sft = 32; sft > 0; sft >>= 1) x += __shfl_xor(x, sft);
      if (lane == 0 && x) atomicAdd(A.totals + (COUNT ? k : 6), x);
    }
  }
}

// Five waves per SIMD: the kernel needs 111 registers per lane (four waves);
// held to 96 it spills 28-60 bytes and runs 5% faster.  The learned kernel
// runs at four waves: with each attempt of a pair recording its own hit it needs
// 127-128 registers and no spills, 4% (persistent) to 11% (one block per chunk)
// faster than at three waves.
template <bool PERSISTENT, bool COUNT>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(5, 5))) k_trace_packet(const WArgs A) {
  packet_body<PERSISTENT, COUNT, false>(A);
}

template <bool PERSISTENT, bool COUNT>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4, 4))) k_trace_packet_learned(const WArgs A) {
  packet_body<PERSISTENT, COUNT, true>(A);
}

// Zeroes a trace control block with a kernel, so a trace recorded into a
// graph holds no memset node.
__global__ void k_zero_words(unsigned int* p, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0u;
}

// Mixed precision, after the trace: every ray whose closest hit was flagged
// (cond(J) above the threshold or predicted t error above t_tol; marked
// prim = -2 - patch by the trace) gets the FP64 correction (correct64_rt:
// guarded FP64 Newton steps from its FP32 (u,v) on FP64 ray planes); t, u, v,
// dpdu, dpdv are replaced and prim restored.  A pass over prim, so the trace kernels carry no FP64 code.
// Flagged rays are scattered, from a few thousand in 8.3M primary rays to
// a tenth of a frame's secondary rays.  The grid is one resident round of
// blocks; each block walks ray ranges of kFixBlock * kFixPer (grid-stride),
// loads a range's patch indices with independent loads and appends its
// flagged rays to an LDS list, and runs the list's FP64 corrections one per
// thread whenever it holds more than a range's worth (and at the end).  A
// correction is a chain of dependent loads and FP64 evaluations: sparse flags
// cost one chain per thread however many ranges a block walks, and dense
// flags spread over all of a block's threads.
constexpr int kFixBlock = 256, kFixPer = 8;
constexpr int kFixRange = kFixBlock * kFixPer, kFixList = 2 * kFixRange;

__device__ __forceinline__ bool fix_one(const WArgs& A, int64_t r) {
  const int pi = -2 - A.prim[r];
  const int pqv = A.pq[pi];
  const int p = pqv & 0xff, q = (pqv >> 8) & 0xff;
  const f3 o{A.ox[r], A.oy[r], A.oz[r]}, d{A.dx[r], A.dy[r], A.dz[r]};
  float u = A.u[r], v = A.v[r], t;
  f3 S, Su, Sv;
  const bool ok = correct64_rt(A.cp + A.cp_off[pi], p, q, o, d, u, v, S, Su, Sv, t);
  if (ok) {
    A.t[r] = t;
    A.u[r] = u;
    A.v[r] = v;
    if (A.dpdu) { A.dpdu[3 * r] = Su.x; A.dpdu[3 * r + 1] = Su.y; A.dpdu[3 * r + 2] = Su.z; }
    if (A.dpdv) { A.dpdv[3 * r] = Sv.x; A.dpdv[3 * r + 1] = Sv.y; A.dpdv[3 * r + 2] = Sv.z; }
  }
  A.prim[r] = pi;
  return ok;
}

// Runs and empties the block's list (called by every thread of the block).
__device__ __forceinline__ void fix_list(const uint32_t* s_list, int* s_n, unsigned long long& done) {
  __syncthreads();
  const int nl = *s_n;
  for (int i = (int)threadIdx.x; i < nl; i += kFixBlock)
    if (fix_one(args_at_use(), (int64_t)s_list[i])) done++;
  __syncthreads();
  if (threadIdx.x == 0) *s_n = 0;
  __syncthreads();
}

__global__ void __launch_bounds__(kFixBlock) k_fix64(const WArgs A) {
  __shared__ uint32_t s_list[kFixList];
  __shared__ int s_n;
  const int tid = threadIdx.x;
  if (tid == 0) s_n = 0;
  __syncthreads();
  const int64_t nranges = (A.n_rays + kFixRange - 1) / kFixRange;
  unsigned long long done = 0;
  for (int64_t g = blockIdx.x; g < nranges; g += gridDim.x) {  // block-uniform loop
    const int64_t base = g * kFixRange;
    int code[kFixPer];
#pragma unroll
    for (int k = 0; k < kFixPer; k++) {
      const int64_t r = base + (int64_t)k * kFixBlock + tid;
      code[k] = r < A.n_rays ? A.prim[r] : -1;
    }
    // s_n <= kFixRange here, so a range's appends stay within kFixList
#pragma unroll
    for (int k = 0; k < kFixPer; k++)
      if (code[k] <= -2) s_list[atomicAdd(&s_n, 1)] = (uint32_t)(base + (int64_t)k * kFixBlock + tid);
    __syncthreads();
    const int nl = s_n;
    __syncthreads();  // every thread has read s_n before any appends to it again
    if (nl > kFixRange) fix_list(s_list, &s_n, done);  // block-uniform
  }
  fix_list(s_list, &s_n, done);
  if (A.totals) {
    for (int sft = 32; sft > 0; sft >>= 1) done += __shfl_xor(done, sft);
    if ((tid & 63) == 0 && done) atomicAdd(A.totals + 8, done);
  }
}

}  // namespace

namespace rtn {

const void* mlp_tiles(const rtn_mlp* net);
int mlp_kind(const rtn_mlp* net);
const float* mlp_bias(const rtn_mlp* net);

namespace {
// One launch: the one-ray form covers n_max rays; the persistent form launches
// blocks_per_cu blocks per compute unit, capped at the blocks the kernel's
// registers and LDS let stay resident (a block that could not start before the
// queue drains would only add a tail).
template <int WB>
void launch_one(void (*kernel)(const WArgs), bool persistent, int blocks_per_cu, int64_t n_max,
                hipStream_t stream, const WArgs& a) {
  const int64_t need = std::max<int64_t>(1, (n_max + WB - 1) / WB);
  unsigned blocks = (unsigned)need;
  if (persistent) {
    int dev = 0, cus = 256, occ = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kernel, WB, 0) != hipSuccess || occ < 1)
      occ = blocks_per_cu;
    const int per_cu = std::max(1, std::min(blocks_per_cu, occ));
    blocks = (unsigned)std::min<int64_t>(need, (int64_t)cus * per_cu);
  }
  hipLaunchKernelGGL(kernel, dim3(blocks), dim3(WB), 0, stream, a);
}
template <int WB, int WSTK, bool P, bool C>
void launch_shape(bool learned, int bpc, int64_t n_max, hipStream_t stream, const WArgs& a) {
  if (!learned) {
    launch_one<WB>(k_trace_wide<WB, WSTK, P, C, false>, P, bpc, n_max, stream, a);
  } else {
    launch_one<WB>(k_trace_wide_learned<WB, WSTK, P, C>, P, bpc, n_max, stream, a);
  }
}
}  // namespace

const LaunchShape kDefaultShape{WB_DEFAULT, WSTK_DEFAULT, 8};
// persistent packet form: resident single-wave blocks per compute unit, capped
// by what the kernel's registers keep resident (20 for the five-wave plain
// kernel, 12 for the three-wave learned kernel): every SIMD slot holds a wave
constexpr int kPacketBlocksPerCU = 32;
const LaunchShape kTunedShapes[] = {{64, 16, 8},  {64, 24, 8},  {64, 32, 8},  {128, 16, 8}, {128, 24, 8},
                                    {128, 32, 8}, {256, 16, 8}, {256, 24, 8}, {256, 32, 8}};
const int kNumTunedShapes = 9;

bool shape_available(const LaunchShape& sh, bool persistent, bool count, bool learned) {
  if (sh.wb == WB_DEFAULT && sh.wstk == WSTK_DEFAULT) return true;
  if (!persistent || count || learned) return false;
  for (int i = 0; i < kNumTunedShapes; i++)
    if (kTunedShapes[i].wb == sh.wb && kTunedShapes[i].wstk == sh.wstk) return true;
  return false;
}

int launch_wide_kernel(const LaunchShape& sh, bool persistent, bool count, bool learned, int64_t n_max,
                       hipStream_t stream, const WArgs& args) {
  RTN_CHECK(shape_available(sh, persistent, count, learned), RTN_EUNSUPPORTED,
            "trace: launch shape %d lanes x %d stack entries is not built for this form", sh.wb, sh.wstk);
  // the persistent queue is a 32-bit counter that every wave may advance by
  // up to 64 once past the end: keep it below 2^32 with room for the grid
  RTN_CHECK(!persistent || n_max <= (int64_t)0xFFFFFFFFu - ((int64_t)1 << 24), RTN_EINVAL,
            "trace: %lld rays exceed the persistent queue's range", (long long)n_max);
  const int bpc = std::max(1, sh.blocks_per_cu);
  if (!persistent) {  // the default shape's stack depth in single-wave blocks
    if (count) launch_shape<WB_ONE_RAY, WSTK_DEFAULT, false, true>(learned, bpc, n_max, stream, args);
    else launch_shape<WB_ONE_RAY, WSTK_DEFAULT, false, false>(learned, bpc, n_max, stream, args);
  } else if (sh.wb == WB_DEFAULT && sh.wstk == WSTK_DEFAULT) {
    if (count) launch_shape<WB_DEFAULT, WSTK_DEFAULT, true, true>(learned, bpc, n_max, stream, args);
    else launch_shape<WB_DEFAULT, WSTK_DEFAULT, true, false>(learned, bpc, n_max, stream, args);
  } else {
#define RTN_SHAPE(W, S) \
  else if (sh.wb == W && sh.wstk == S) launch_one<W>(k_trace_wide<W, S, true, false, false>, true, bpc, n_max, stream, args);
    if (false) {
    }
    RTN_SHAPE(64, 16) RTN_SHAPE(64, 24) RTN_SHAPE(64, 32) RTN_SHAPE(128, 16) RTN_SHAPE(128, 24) RTN_SHAPE(128, 32)
    RTN_SHAPE(256, 16) RTN_SHAPE(256, 24) RTN_SHAPE(256, 32)
#undef RTN_SHAPE
  }
  RTN_HIP(hipGetLastError());
  return launch_fix64(n_max, stream, args);
}

int fill_wide_args(const rtn_scene* s, const float* const rays[6], int64_t n, const rtn_trace_params* prm,
                   const rtn_hits* hits, unsigned long long* d_totals, const PairOut* pairs, WArgs& A) {
  const bool learned = prm->seed_mode == RTN_SEED_LEARNED;
  RTN_CHECK(prm->seed_mode == RTN_SEED_PLAIN || learned, RTN_EINVAL, "rtn_trace: unknown seed mode %d",
            prm->seed_mode);
  RTN_CHECK(!learned || prm->seed_net, RTN_EINVAL, "rtn_trace: the learned seed needs seed_net");
  RTN_CHECK(prm->learned_iters >= 0 && prm->learned_iters <= 64, RTN_EINVAL, "rtn_trace: learned_iters outside 0..64");
  RTN_CHECK(prm->fallback_iters >= 0 && prm->fallback_iters <= 64, RTN_EINVAL,
            "rtn_trace: fallback_iters outside 0..64");
  RTN_CHECK(prm->seed_share == 0 || prm->seed_share == 1, RTN_EINVAL, "rtn_trace: seed_share is 0 or 1");
  RTN_CHECK(prm->stall_exit == 0 || prm->stall_exit == 1, RTN_EINVAL, "rtn_trace: stall_exit is 0 or 1");
  RTN_CHECK(prm->precision != RTN_PREC_MIXED || (hits->t && hits->u && hits->v), RTN_EINVAL,
            "rtn_trace: mixed precision needs the t, u and v outputs");
  RTN_CHECK(prm->flat_threshold <= 0.f, RTN_EUNSUPPORTED, "rtn_trace: flat-skip runs in the exact form only");
  RTN_CHECK(prm->hit_gate >= 0.f && prm->hit_gate <= 1.f && prm->hit_skip >= 0.f && prm->hit_skip <= 1.f, RTN_EINVAL,
            "rtn_trace: hit_gate and hit_skip lie in [0, 1]");
  RTN_CHECK(prm->learned_min_lanes >= 0 && prm->learned_min_lanes <= 64, RTN_EINVAL,
            "rtn_trace: learned_min_lanes outside 0..64");
  A = WArgs{};
  A.nodes = s->wide.ptr;
  A.cp = s->cp.ptr;
  A.cp_off = s->cp_off.ptr;
  A.pq = s->pq.ptr;
  A.corners = s->corners.ptr;
  A.obb = s->obb.ptr;
  A.seed64 = s->seed64.ptr;
  A.ox = rays[0]; A.oy = rays[1]; A.oz = rays[2]; A.dx = rays[3]; A.dy = rays[4]; A.dz = rays[5];
  A.n_rays = n;
  A.max_iter = prm->max_iter;
  A.thr_det = prm->thr_det;
  A.thr_res = prm->thr_res;
  A.seed_base = prm->seed_base;
  A.hull_cull = prm->hull_cull;
  A.cull_margin = 2.0f * prm->thr_res + 1e-5f;
  A.mixed = prm->precision == RTN_PREC_MIXED;
  A.cond_thr = prm->cond_thr;
  A.t_tol = prm->t_tol > 0.f ? prm->t_tol : 2.5e-6f;
  A.any_hit = prm->any_hit;
  A.stall_exit = prm->stall_exit;
  A.refill_min = prm->refill_min > 0 ? prm->refill_min : 64;
  if (learned) {
    A.mlp_tiles = reinterpret_cast<const mlp::bf16x8*>(mlp_tiles(prm->seed_net));
    A.mlp_bias = mlp_bias(prm->seed_net);
    A.bounds = s->bounds.ptr;
    A.learned_iters = prm->learned_iters > 0 ? prm->learned_iters : prm->max_iter;
    A.fallback_iters = prm->fallback_iters > 0 ? prm->fallback_iters : prm->max_iter;
    A.seed_share = prm->seed_share;
    A.mlp_kind = mlp_kind(prm->seed_net);
    A.hit_gate = prm->hit_gate;
    A.hit_skip = prm->hit_skip;
    A.learned_min_lanes = prm->learned_min_lanes;
  }
  if (pairs) {
    A.pair_ray = pairs->ray;
    A.pair_patch = pairs->patch;
    A.pair_uv = pairs->uv;
    A.pair_count = pairs->count;
    A.pair_cap = pairs->cap;
    A.pair_all = pairs->all;
  }
  A.prim = hits->prim; A.t = hits->t; A.u = hits->u; A.v = hits->v; A.dpdu = hits->dpdu; A.dpdv = hits->dpdv;
  A.skip_prim = prm->skip_patch;
  A.skip_parent = nullptr;
  A.counters = prm->count ? hits->counters : nullptr;
  A.totals = d_totals;
  return RTN_OK;
}

int trace_wide_scene(const rtn_scene* s, const float* const rays[6], int64_t n, const rtn_trace_params* prm,
                     const rtn_hits* hits, unsigned long long* d_totals, hipStream_t stream,
                     const PairOut* pairs) {
  WArgs A;
  RTN_TRY(fill_wide_args(s, rays, n, prm, hits, d_totals, pairs, A));
  const bool persistent = prm->mode == RTN_MODE_PERSISTENT;
  A.work = s->ctl.ptr->work;  // zeroed by the caller's ctl_reset
  return launch_wide_kernel(kDefaultShape, persistent, prm->count != 0, prm->seed_mode == RTN_SEED_LEARNED, n,
                            stream, A);
}

int ctl_reset(TraceCtl* ctl, hipStream_t stream) {
  hipLaunchKernelGGL(k_zero_words, dim3(1), dim3(256), 0, stream, reinterpret_cast<unsigned int*>(ctl),
                     (int)(sizeof(TraceCtl) / sizeof(unsigned int)));
  RTN_HIP(hipGetLastError());
  return RTN_OK;
}

int launch_fix64(int64_t n_max, hipStream_t stream, const WArgs& args) {
  if (!args.mixed || n_max <= 0) return RTN_OK;
  // one resident round (at most one block per range)
  static const int resident = [] {  // blocks per compute unit x compute units, queried once
    int dev = 0, cus = 256, occ = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_fix64, kFixBlock, 0) != hipSuccess || occ < 1) occ = 1;
    return occ * cus;
  }();
  const unsigned blocks = (unsigned)std::max<int64_t>(1, std::min<int64_t>((n_max + kFixRange - 1) / kFixRange, resident));
  hipLaunchKernelGGL(k_fix64, dim3(blocks), dim3(kFixBlock), 0, stream, args);
  RTN_HIP(hipGetLastError());
  return RTN_OK;
}

int launch_packet_kernel(bool persistent, bool count, bool learned, int64_t n_max, hipStream_t stream,
                         const WArgs& args) {
  RTN_CHECK(!persistent || n_max <= (int64_t)0xFFFFFFFFu - ((int64_t)1 << 24), RTN_EINVAL,
            "trace: %lld rays exceed the persistent queue's range", (long long)n_max);
  RTN_CHECK(args.bounds, RTN_EINVAL, "trace: the packet form needs the patch boxes");
  void (*kernel)(const WArgs);
  if (persistent) {
    kernel = learned ? (count ? k_trace_packet_learned<true, true> : k_trace_packet_learned<true, false>)
                     : (count ? k_trace_packet<true, true> : k_trace_packet<true, false>);
  } else {
    kernel = learned ? (count ? k_trace_packet_learned<false, true> : k_trace_packet_learned<false, false>)
                     : (count ? k_trace_packet<false, true> : k_trace_packet<false, false>);
  }
  launch_one<64>(kernel, persistent, kPacketBlocksPerCU, n_max, stream, args);
  RTN_HIP(hipGetLastError());
  return launch_fix64(n_max, stream, args);
}

int trace_packet_scene(const rtn_scene* s, const float* const rays[6], int64_t n, const rtn_trace_params* prm,
                       const rtn_hits* hits, unsigned long long* d_totals, hipStream_t stream) {
  const bool persistent = prm->mode == RTN_MODE_PERSISTENT_PACKET;
  WArgs A;
  RTN_TRY(fill_wide_args(s, rays, n, prm, hits, d_totals, nullptr, A));
  A.bounds = s->bounds.ptr;
  A.work = s->ctl.ptr->work;  // zeroed by the caller's ctl_reset
  return launch_packet_kernel(persistent, prm->count != 0, prm->seed_mode == RTN_SEED_LEARNED, n, stream, A);
}

}  // namespace rtn
